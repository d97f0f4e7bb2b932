// internal.h -- host-side structures shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dtwax.h"
#include "kernels.h"

namespace dtwax {

int set_error(int code, const std::string& msg);

#define DTWAX_HIP_TRY(expr)                                                                     \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return ::dtwax::set_error(DTWAX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A run of work items sharing one segment width (one persistent launch).
struct LaunchGroup {
  int S = 32;
  int32_t item0 = 0;   // first item in the item array
  int32_t n_items = 0;
  int grid = 1;
  float renorm_k = 0.f;
  bool renorm_split = false;  // KParams::renorm_split
  int64_t scratch_wg0 = 0;  // first band-scratch row of this group
  bool pack = false;        // packed items (sdtw_f16_pack_kernel)
  int32_t rows = 1;         // band scratch rows per workgroup
  hipStream_t stream = nullptr;  // groups after the first run on their own stream
  hipEvent_t done = nullptr;     // recorded on that stream after the launch
};

// One scoring pass: reset keys and counters, one persistent sDTW launch per
// group (in order, on `st`), finalise `n_reads` results.  Without groups the
// fields S/items/n_items/grid/renorm_k describe the single launch.  ev0/ev1
// (optional) bracket the sDTW launches.
struct ScoreLaunch {
  int precision = DTWAX_F16;
  int S = 32;
  int nstr = 2;
  int grid = 1;
  const float* queries = nullptr;
  const int64_t* q_off = nullptr;
  const int32_t* q_len = nullptr;
  const WorkItem* items = nullptr;
  int32_t n_items = 0;
  int32_t n_reads = 0;
  int32_t* counter = nullptr;
  unsigned long long* keys = nullptr;
  void* scratch = nullptr;
  int64_t scratch_stride = 0;
  void* state = nullptr;
  int64_t row_stride = 0;
  float renorm_k = 0.f;
  bool m3 = false;  // three moves
  // binary16 sweeps may order a SIMD's waves by progress (KParams::
  // prio_by_progress) when the launch has more waves than SIMDs; the Read
  // Until engine turns this off: its batches share SIMDs with each other, and
  // a later batch's high-priority waves held back earlier ones (decision
  // latency p99 6.0-6.5 -> 6.9-8.0 ms on a 512-channel flow cell)
  bool prio_ok = true;
  DeviceResult* out = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const PackLane* pack_lanes = nullptr;  // packed groups' lane and read tables
  const PackRead* pack_reads = nullptr;
  const LaunchGroup* groups = nullptr;  // counter[g] for group g
  int n_groups = 0;
  hipEvent_t fork = nullptr;  // recorded on the caller's stream before the groups start
  // Reference-layout batches (dtwax_distances*, launch_sdtw_fixed): a
  // reference other than the context's (column 0 of a padded buffer, n_cols
  // columns), fixed-stride queries, the seam deviations, the tiled last row,
  // and results in the reference's dists layout instead of `out`.
  bool fixed = false;
  const void* ref = nullptr;
  int64_t n_cols = 0;
  int32_t q_stride = 0;
  bool quirks = false;
  void* last_row = nullptr;
  int64_t lr_stride = 0;
  int32_t lr_ntiles = 0, lr_last = 0;
  void* dists = nullptr;
  bool dists_h2 = false;
  const uint32_t* ru_renorm = nullptr;  // Read Until DTWAX_F16_RENORM: {K, k0} per read
};

}  // namespace dtwax

struct dtwax_ctx {
  int device = 0;
  int n_cu = 0;
  int clock_khz = 0;
  std::string name;
  // reference
  int32_t n_strands = 0;
  int64_t n_cols = 0;
  void* ref16_alloc = nullptr;  // h2 per column, COL_PADL + n + COL_PADR
  void* ref32_alloc = nullptr;  // n_strands floats per column, same padding
  // Buffers of the reference-layout entry points (dtwax_distances for the
  // FP16 build, dtwax_distances_f32 for the FP32 build), apart from the
  // uploaded reference: the untiled reference of each (padded like it),
  // allocated on first use, and one work counter + 2 keys per read, grown
  // to the largest batch.
  void* compat_ref[2] = {nullptr, nullptr};  // [0] half2, [1] float
  int32_t* compat_counter = nullptr;
  unsigned long long* compat_keys = nullptr;
  int32_t compat_cap = 0;
};

// A plan: work items (read x reference chunk) in decreasing cost order and
// the buffers one launch needs.
struct dtwax_plan {
  dtwax_ctx* ctx = nullptr;
  int32_t n_reads = 0;
  int precision = DTWAX_F16;
  bool m3 = false;  // three moves (whole-strand items)
  int S = 32;  // 0 when the reads are split over several segment widths
  std::vector<dtwax::LaunchGroup> groups;  // launched concurrently, one stream each
  hipEvent_t fork = nullptr;
  int nstr = 2;
  int32_t chunk = 0;
  int grid = 0;
  int32_t n_items = 0;
  int64_t cells = 0, computed = 0;
  float renorm_k = 0.f;
  dtwax::WorkItem* d_items = nullptr;
  dtwax::PackLane* d_pack_lanes = nullptr;
  dtwax::PackRead* d_pack_reads = nullptr;
  int32_t n_packs = 0, n_packed_reads = 0;
  int32_t* d_counter = nullptr;
  unsigned long long* d_keys = nullptr;
  void* d_scratch = nullptr;
  int64_t scratch_stride = 0;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed_once = false;
};

namespace dtwax {
int run_score_launch(dtwax_ctx* c, const ScoreLaunch& L, hipStream_t st);
int occupancy(int precision, int S, int nstr, bool m3);
// dtwax_plan_create over reference columns [0, col_limit) only.  dry: the
// plan's layout only (widths, chunk, items, grid, cell counts); nothing is
// allocated on the device and the plan cannot run.
int plan_build(dtwax_ctx* c, const int32_t* h_lens, int32_t n_reads, const dtwax_score_opts* opts,
               int64_t col_limit, dtwax_plan** out, bool dry = false);
}  // namespace dtwax
