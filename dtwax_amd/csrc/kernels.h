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
  // F16_RENORM at S = 16 (not Read Until): nonzero runs the split schedule,
  // the first Q - renorm_split_rows(Q) rows in bands of 64 x 8 rows from
  // drift renorm_k, the rest in bands of 64 x 16 rows at twice the drift
  int32_t renorm_split;
  const PackLane* pack_lanes;  // 64 per pack (packed launches)
  const PackRead* pack_reads;
  int32_t scratch_rows;        // band scratch rows per workgroup (packed launches; else 1)
  // binary16 sweeps: nonzero lowers each wave's issue priority as it moves
  // through a band (set for every group of a launch when one of them has
  // more waves than the GPU has SIMDs)
  int32_t prio_by_progress;
  // Fixed-geometry kernels (launch_sdtw_fixed): read i is q_stride samples
  // from i * q_stride (q_off, q_len and items unused).
  int32_t q_stride;
  // Fixed-geometry kernels: when not null, the final band's last row also goes to
  // last_row[read * lr_stride + tiled(col)], the reference's tiled layout:
  // tile t = col / TILE_W at t * TILE_W (the last of lr_ntiles tiles at
  // lr_last), column 26 j + i of a tile at j + 32 i.
  void* last_row;
  int64_t lr_stride;
  int32_t lr_ntiles;
  int32_t lr_last;
  // Read Until items of DTWAX_F16_RENORM: per read of the batch two half2
  // words {K, k0}: K = the per-lane drift subtracted at every lane hop of
  // every band, k0 = what lane 0 also subtracts from a resumed top row (the
  // stored row's minimum).  The keys then hold the lowered scores; the
  // engine adds the offsets back (per-slot, in binary64).
  const uint32_t* ru_renorm;
};

// The reference's tile geometry (src/include/common.hpp:63-66, 87):
// SEGMENT_SIZE 26 columns per lane, WARP_SIZE 32 lanes, tiles of 832 columns.
constexpr int REF_SEG = 26;
constexpr int REF_WARP = 32;
constexpr int TILE_W = REF_SEG * REF_WARP;

// m3: three moves (the left move D[r][c-1] too; DTWAX_F16 and DTWAX_F32 only)
// pack: the packed kernel (DTWAX_F16 / DTWAX_F16_FUSED, S = 16, 20, 25, 28 or 32, two moves)
hipError_t launch_sdtw(const KParams& p, int precision, int S, int nstr, int grid, bool m3,
                       hipStream_t st, bool pack = false);
// The fixed-geometry kernels of the reference-layout entry points: S = 32,
// two moves, every read q_stride <= 2048 rows (one band), items implicit
// (item i = read i over [0, n_cols) from x = 0); DTWAX_F16 reads half2 query
// words and may apply the reference kernel's seam deviations (quirks);
// DTWAX_F32 is one strand of float queries.
hipError_t launch_sdtw_fixed(const KParams& p, int precision, int grid, bool quirks, hipStream_t st);
int occupancy_fixed(int precision, bool quirks);
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

// keys -> the reference's output layout: one half2 word {fwd, rev} per read
// (h2 = true, the FP16 build) or one float per read (the FP32 build, strand 0)
hipError_t launch_finalize_dists(const unsigned long long* keys, int32_t n_reads, bool h2, void* dists,
                                 hipStream_t st);

// reference-compat helpers (compat_kernels.hip)
// tiled 32-bit words (half2 or float; src/main.cu:92-109) -> flat column order
hipError_t launch_untile_reference(const uint32_t* tiled, int32_t n_tiles, uint32_t* flat, hipStream_t st);
// The reference kernel's schedule itself, one wave per read draining at every
// tile end (QUERY_LEN 2048, PREFIX_LEN 1024, tiles of 832 columns), with its
// seam deviations: ref = flat columns (half2 {fwd, rev} words when h2, else
// floats), queries = num_entries x 2048 half2 words {q, q} (h2) or floats;
// dists and last_row (optional) in the reference's layouts.
hipError_t launch_compat_tiles(bool h2, const uint32_t* ref, const void* queries, int32_t n_reads,
                               int32_t n_tiles, void* dists, void* last_row, int64_t lr_stride,
                               int32_t lr_last, hipStream_t st);

}  // namespace dtwax
