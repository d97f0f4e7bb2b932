// compat_kernels.hip -- layout adapters for the reference-compatible entry
// point dtwax_distances (mirror of FullDTW::distances, DTW.hpp:285-295).
//
//   untile:   the reference's tiled half2 squiggle (src/main.cu:92-109:
//             h[l*TILE + i*WARP + j] = flat[l*TILE + j*SEG + i]) back to
//             flat column order, into the padded streaming layout.
//   queries:  half2 {q, q} (main.cu:170-175) -> float q (exact).
//   results:  {fwd, rev} float scores (binary16 values) -> half2.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace dtwax {

__global__ void untile_kernel(const uint32_t* __restrict__ tiled, int32_t seg, int32_t warp,
                              int32_t n_cols, uint32_t* __restrict__ flat) {
  int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const int32_t tile = seg * warp;
  const int32_t l = c / tile, rem = c % tile;
  const int32_t j = rem / seg, i = rem % seg;
  flat[c] = tiled[(int64_t)l * tile + (int64_t)i * warp + j];
}

__global__ void q_h2_to_f32_kernel(const uint32_t* __restrict__ q, int64_t n, float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint16_t lo = (uint16_t)(q[i] & 0xffffu);
  out[i] = (float)__builtin_bit_cast(_Float16, lo);
}

__global__ void res_to_h2_kernel(const DeviceResult* __restrict__ r, int32_t n, uint32_t* __restrict__ d) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  _Float16 a = (_Float16)r[i].score[0], b = (_Float16)r[i].score[1];
  d[i] = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

hipError_t launch_untile_reference(const uint32_t* tiled, int32_t ref_len, int32_t seg, int32_t warp,
                                   int32_t n_tiles, uint32_t* untiled_h2, hipStream_t st) {
  (void)ref_len;
  int32_t n = n_tiles * seg * warp;
  hipLaunchKernelGGL(untile_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tiled, seg, warp, n,
                     untiled_h2);
  return hipGetLastError();
}

hipError_t launch_half2_queries_to_float(const uint32_t* q_h2, int64_t n, float* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(q_h2_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q_h2, n,
                     out);
  return hipGetLastError();
}

hipError_t launch_results_to_half2(const DeviceResult* res, int32_t n, uint32_t* dists, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(res_to_h2_kernel, dim3((n + 255) / 256), dim3(256), 0, st, res, n, dists);
  return hipGetLastError();
}

}  // namespace dtwax
