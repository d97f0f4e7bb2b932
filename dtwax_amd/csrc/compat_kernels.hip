// compat_kernels.hip -- kernels of the reference-layout entry points
// dtwax_distances / dtwax_distances_f32 (mirrors of FullDTW::distances,
// src/include/DTW.hpp:285-295, in the FP16 and FP32 builds).
//
//   untile:   the reference's tiled squiggle (src/main.cu:92-109:
//             h[l*TILE + i*WARP + j] = flat[l*TILE + j*SEG + i]) back to flat
//             column order, 32-bit words (half2 {fwd, rev} or float).
//   tiles:    the reference kernel's own schedule with its seam deviations,
//             one wave per read, as an independent second implementation of
//             the deviations (the persistent kernel applies them in its
//             two-column sweep, sdtw_kernels.hip).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace dtwax {

__global__ void untile_kernel(const uint32_t* __restrict__ tiled, int32_t n_cols, uint32_t* __restrict__ flat) {
  int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const int32_t l = c / TILE_W, rem = c % TILE_W;
  const int32_t j = rem / REF_SEG, i = rem % REF_SEG;
  flat[c] = tiled[(int64_t)l * TILE_W + (int64_t)i * REF_WARP + j];
}

typedef _Float16 half_t;
typedef half_t h2 __attribute__((ext_vector_type(2)));

template <class T>
__device__ __forceinline__ T shfl_w(T v, int src) {
  return __builtin_bit_cast(T, __shfl(__builtin_bit_cast(int, v), src, 64));
}
template <class T> __device__ __forceinline__ T splat(float v);
template <> __device__ __forceinline__ h2 splat<h2>(float v) { return (h2){(half_t)v, (half_t)v}; }
template <> __device__ __forceinline__ float splat<float>(float v) { return v; }
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return __builtin_elementwise_min(a, b); }

// The reference kernel's deviations from the clean recurrence (QUERY_LEN
// 2048, PREFIX_LEN 1024, SEGMENT 26, WARP 32; tiles of 832 columns),
// reproduced on top of the clean per-op recurrence (binary16 half2 for the
// FP16 build, binary32 for the FP32 build, common.hpp:32-57):
//   Q1  global row 1024 (first row of the second 1024-row pass) takes +inf
//       as its diagonal at c = 832k (k >= 1) and c = 832k + 26 (k >= 0):
//       lane 0's diagonal is reset at every tile start (DTW.cu:315, 442) and
//       lane 1's first diagonal is consumed before it is filled
//       (DTW.cu:179-180, 238-240);
//   Q2  at c = 832k (k >= 2) rows 962..992 of each pass take D[r + 30][c - 1]
//       as their diagonal: middle tiles flush the last column to entry
//       wave - 63 instead of wave - 32 (DTW.cu:385-388), so entries 961..991
//       hold rows 992..1022.
// One wave per read, 64 lanes x 32 rows = 2048 rows, lane L at column
// 832t + T - L at step T of tile t.  The wave drains at every tile end like
// the reference, so at a tile's first column the previous tile's last column
// of every row is still in the lanes' registers (the Q2 source rows live in
// lanes 31 and 63, which start the tile after lanes 30 and 62 use them).
// Inactive lanes keep their registers through selects and the boundary
// values move by ds_bpermute: about half the persistent kernel's speed.
template <class T>
__global__ __launch_bounds__(64) void sdtw_compat_tiles_kernel(const uint32_t* __restrict__ ref,
                                                               const uint32_t* __restrict__ queries,
                                                               int32_t n_tiles, uint32_t* __restrict__ dists,
                                                               uint32_t* __restrict__ last_row, int64_t lr_stride,
                                                               int32_t lr_last) {
  constexpr int S = 32;
  const int lane = threadIdx.x;
  const uint32_t* q = queries + (int64_t)blockIdx.x * 2048;
  const T INF = splat<T>(__builtin_inff());
  const T ZERO = splat<T>(0.0f);
  const T ONE = splat<T>(1.0f);
  T Qn[S], D[S], ov[S];
  for (int i = 0; i < S; i++) {
    Qn[i] = -__builtin_bit_cast(T, q[lane * S + i]);  // {-q_fwd, -q_rev} or -q
    D[i] = INF;  // column -1
    ov[i] = INF;
  }
  uint32_t* lrow = last_row ? last_row + (int64_t)blockIdx.x * lr_stride : nullptr;
  T tin = INF, dg = INF;
  T best = INF;
  for (int32_t t = 0; t < n_tiles; t++) {
    // Q2 sources: lanes 30 and 62 take rows 992.. / 2016.. of the previous
    // tile's last column from lanes 31 and 63
    for (int i = 2; i < S; i++) ov[i] = shfl_w(D[i - 2], lane + 1 < 64 ? lane + 1 : lane);
    const bool q2_tile = t >= 2;
    const int64_t tile_base = t < n_tiles - 1 ? (int64_t)t * TILE_W : lr_last;
    for (int32_t step = 0; step < TILE_W + 63; step++) {
      const int32_t cl = step - lane;
      const bool active = cl >= 0 && cl < TILE_W;
      const int32_t c = t * TILE_W + (active ? cl : 0);
      const T s = __builtin_bit_cast(T, ref[c]);
      T tp = (lane == 0) ? ZERO : tin;
      T dd = (lane == 0) ? ZERO : dg;
      if (lane == 32 && ((cl == 0 && t >= 1) || cl == 26)) dd = INF;  // Q1 (row 1024)
      const bool q2_col = q2_tile && cl == 0;
      if (q2_col && (lane == 31 || lane == 63)) dd = D[30];          // Q2 rows 992 / 2016
      for (int i = 0; i < S; i++) {
        const T tt = __builtin_elementwise_fma(s, ONE, Qn[i]);
        const T c2 = tt * tt;
        const T m = vmin(tp, dd);
        const T v = c2 + m;
        const bool q2_row = q2_col && (lane == 30 || lane == 62) && i >= 1 && i + 1 < S;
        dd = q2_row ? ov[i + 1] : D[i];  // diagonal of row i + 1
        D[i] = active ? v : D[i];
        tp = active ? v : tp;
      }
      if (lane == 63 && active) {
        best = vmin(best, D[S - 1]);
        if (lrow) {
          const int32_t j = cl / REF_SEG, i = cl - j * REF_SEG;
          lrow[tile_base + j + REF_WARP * i] = __builtin_bit_cast(uint32_t, D[S - 1]);
        }
      }
      // boundary: lane L's next top input is lane L-1's last row now
      dg = tin;
      tin = shfl_w(D[S - 1], lane > 0 ? lane - 1 : 0);
    }
  }
  if (lane == 63) dists[blockIdx.x] = __builtin_bit_cast(uint32_t, best);
}

hipError_t launch_compat_tiles(bool half2_build, const uint32_t* ref, const void* queries, int32_t n_reads,
                               int32_t n_tiles, void* dists, void* last_row, int64_t lr_stride,
                               int32_t lr_last, hipStream_t st) {
  if (n_reads <= 0) return hipSuccess;
  if (half2_build)
    hipLaunchKernelGGL(sdtw_compat_tiles_kernel<h2>, dim3(n_reads), dim3(64), 0, st, ref,
                       (const uint32_t*)queries, n_tiles, (uint32_t*)dists, (uint32_t*)last_row, lr_stride,
                       lr_last);
  else
    hipLaunchKernelGGL(sdtw_compat_tiles_kernel<float>, dim3(n_reads), dim3(64), 0, st, ref,
                       (const uint32_t*)queries, n_tiles, (uint32_t*)dists, (uint32_t*)last_row, lr_stride,
                       lr_last);
  return hipGetLastError();
}

hipError_t launch_untile_reference(const uint32_t* tiled, int32_t n_tiles, uint32_t* flat, hipStream_t st) {
  const int32_t n = n_tiles * TILE_W;
  hipLaunchKernelGGL(untile_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tiled, n, flat);
  return hipGetLastError();
}

}  // namespace dtwax
