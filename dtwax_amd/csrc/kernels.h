// kernels.h -- device-side structures shared by the HIP kernels and the host
// launcher (internal; the public C ABI is include/dtwax.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dtwax.h"

namespace dtwax {

// Padding (in columns) around every device-resident row: the reference, the
// band scratch rows and the Read Until state rows.  The sweep reads columns
// [x - 64, c1 + 253].
constexpr int COL_PADL = 64;
constexpr int COL_PADR = 384;
constexpr int SCRATCH_PADL = COL_PADL;

// One work item: read `read` against reference columns [x, c1), counting the
// min only over [c0, c1) (x = c0 - (Q-1) clamped at 0: columns [x, c0) are
// the overlap that makes a chunk exact under the no-deletion recurrence).
struct WorkItem {
  int32_t read;
  int32_t slot;  // Read Until state slot, else 0
  int32_t x;
  int32_t c0;
  int32_t c1;
  int32_t flags;  // ITEM_* below (Read Until only)
  int32_t cur;    // Read Until: which of the slot's two rows holds the current last row
  int32_t reserved;
};

// Read Until item flags: resume from the slot's stored row (else free start);
// store the new last row into the slot's other row.
constexpr int32_t ITEM_RESUME = 1;
constexpr int32_t ITEM_STORE = 2;
// A packed item: `read` is a pack, whose shared band holds the first
// (partial) band of several reads on consecutive lanes; `slot` is its first
// PackRead, `cur` the number of PackReads (reads with full bands after the
// shared band), `reserved` the shared band's history flush period in steps
// (a power of two, at most the shortest segment's lane count).
constexpr int32_t ITEM_PACK = 4;

// One lane of a pack's shared band.  A read's rows [0, rp) fill lanes
// [start, end] with the rows padded at the front to a lane boundary
// (virtual rows, zero cost).
struct PackLane {
  int32_t read;  // -1: idle lane
  int32_t r0;    // read row of the lane's first row (< 0: virtual rows first)
  int32_t info;  // segment end lane (bits 0-7) | PL_* flags
  int32_t row;   // band scratch row of the read when PL_HAND, else -1
};
constexpr int32_t PL_START = 256;  // first lane of a read: top input is the free start
constexpr int32_t PL_FINAL = 512;  // the read ends in this band: its min is scored here
constexpr int32_t PL_HAND = 1024;  // the read continues: its last row goes to `row`

// A packed read with full bands after the shared band: rows rp + b * 64S ...
struct PackRead {
  int32_t read;
  int32_t nfull;  // full bands after the shared band
  int32_t rp;     // its rows in the shared band
  int32_t row;    // band scratch row (within the wave's rows)
};

struct DeviceResult {
  float score[2];
  int32_t argmin[2];
};

struct KParams {
  const void* ref;  // column 0 of the padded reference (h2 or NSTR floats per column)
  int32_t n_cols;
  int32_t n_items;
  const float* queries;
  const int64_t* q_off;
  const int32_t* q_len;
  const WorkItem* items;
  int32_t* counter;
  unsigned long long* keys;  // 2 per read: (ordered score bits << 32) | column
  void* scratch;             // per workgroup band rows
  int64_t scratch_stride;    // columns per workgroup (includes the pads)
  void* state;               // Read Until: column 0 of slot 0's row 0; slot s row k
                             // starts (2*s + k) * row_stride columns further
  int64_t row_stride;        // columns per state row (includes the pads)
  float renorm_k;            // F16_RENORM per-lane drift constant (S rows' worth)
  const PackLane* pack_lanes;  // 64 per pack (packed launches)
  const PackRead* pack_reads;
  int32_t scratch_rows;        // band scratch rows per workgroup (packed launches; else 1)
};

// m3: three moves (the left move D[r][c-1] too; DTWAX_F16 and DTWAX_F32 only)
// pack: the packed kernel (DTWAX_F16 / DTWAX_F16_FUSED, S = 16, 20, 25, 28 or 32, two moves)
hipError_t launch_sdtw(const KParams& p, int precision, int S, int nstr, int grid, bool m3,
                       hipStream_t st, bool pack = false);
// zero n_counters work counters (n_counters <= 256) and set n_keys keys to ~0
hipError_t launch_reset(int32_t* counters, int32_t n_counters, unsigned long long* keys, int64_t n_keys,
                        hipStream_t st);
hipError_t launch_finalize(const unsigned long long* keys, int32_t n_reads, int32_t n_strands,
                           DeviceResult* out, hipStream_t st);
hipError_t launch_selftest_lanes(int32_t* d_out, hipStream_t st);
int occupancy_f16(int precision, int S, bool m3);
int occupancy_f16_pack(int precision, int S);
int occupancy_f32(int S, int nstr, bool m3);

// normalisation (normalize_kernels.hip)
hipError_t launch_normalize(const float* raw, const int64_t* off, const int32_t* len, int32_t n_reads,
                            int mode, double eps, float* out, double* stats, hipStream_t st);

// Read Until chunk normalisation: first chunks fix (centre, inv) per slot,
// later chunks apply the slot's values.
hipError_t launch_ru_normalize(const float* raw, int32_t n_chunks, int32_t chunk_len, const int32_t* slot,
                               const int32_t* first, int mode, double eps, double* slot_stats,
                               float* out, hipStream_t st);

// fill n 32-bit words with `value` (a kernel, so it may be captured in a graph)
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t st);

// reference-compat helpers (compat_kernels.hip)
hipError_t launch_untile_reference(const uint32_t* tiled, int32_t ref_len, int32_t seg, int32_t warp,
                                   int32_t n_tiles, uint32_t* untiled_h2, hipStream_t st);
hipError_t launch_half2_queries_to_float(const uint32_t* q_h2, int64_t n, float* out, hipStream_t st);
hipError_t launch_results_to_half2(const DeviceResult* res, int32_t n, uint32_t* dists, hipStream_t st);
// binary16 scores with the reference kernel's seam deviations (QUERY_LEN 2048,
// PREFIX_LEN 1024, tiles of 832 columns): ref_h2 = flat {fwd, rev} columns
hipError_t launch_compat_quirks(const uint32_t* ref_h2, const float* queries, int32_t n_reads, int32_t n_tiles,
                                DeviceResult* out, hipStream_t st);

}  // namespace dtwax
