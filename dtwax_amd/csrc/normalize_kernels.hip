// normalize_kernels.hip -- per-read z-normalisation with FP64 statistics.
//
// One 256-thread workgroup per read.
//   MEANSTD: y = (x - mean) / sqrt(var + eps), biased variance (the
//            arithmetic of cudnnBatchNormalizationForwardTraining in SPATIAL
//            mode that src/include/normalizer.cu:86-100 calls, eps 1e-4
//            there), two-pass FP64 sums.
//   MEDMAD:  y = (x - median) / (1.4826 * MAD); median and MAD are exact
//            order statistics (mean of the two middle ones for even n),
//            found by radix selection on order-preserving integer keys
//            (32-bit keys of the float samples, 64-bit keys of the FP64
//            absolute deviations), 8 bits per pass with an LDS histogram.
//   NONE:    copy.
// Results are written as float; stats (optional) receive (centre, scale).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace dtwax {

constexpr int NT = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; i++) t += sh[i];
    sh[NT / 64] = t;
  }
  __syncthreads();
  return sh[NT / 64];
}

__device__ __forceinline__ uint32_t fkey(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __builtin_bit_cast(float, u);
}
__device__ __forceinline__ uint64_t dkey(double d) {
  uint64_t u = __builtin_bit_cast(uint64_t, d);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(uint64_t k) {
  uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, u);
}

// k-th smallest (0-based) key over n elements, keys computed by `keyf(i)`.
template <int BITS, class KeyF>
__device__ uint64_t radix_select(int32_t n, int64_t k, KeyF keyf, uint32_t* hist, uint64_t* shv) {
  uint64_t prefix = 0, mask = 0;
  for (int shift = BITS - 8; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += NT) hist[i] = 0;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += NT) {
      uint64_t key = keyf(i);
      if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 0xff], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t acc = 0;
      int bin = 0;
      for (; bin < 256; bin++) {
        if (acc + (int64_t)hist[bin] > k) break;
        acc += hist[bin];
      }
      shv[0] = prefix | ((uint64_t)bin << shift);
      shv[1] = (uint64_t)(k - acc);
    }
    __syncthreads();
    prefix = shv[0];
    k = (int64_t)shv[1];
    mask |= (uint64_t)0xff << shift;
    __syncthreads();
  }
  return prefix;
}

// Statistics of one read: centre, scale and the multiplier inv applied as
// y = (float)((x - centre) * inv).  NONE gives centre 0, inv 1 (a copy).
__device__ void read_stats(const float* __restrict__ x, int32_t n, int mode, double eps, double* sh,
                           uint32_t* hist, uint64_t* shv, double& centre, double& scale,
                           double& inv) {
  centre = 0.0;
  scale = 1.0;
  inv = 1.0;
  if (mode == DTWAX_NORM_MEANSTD) {
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += NT) s += (double)x[i];
    const double mean = block_sum(s, sh) / (double)n;
    double v = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += NT) {
      double d = (double)x[i] - mean;
      v += d * d;
    }
    const double var = block_sum(v, sh) / (double)n;
    centre = mean;
    scale = sqrt(var + eps);
    inv = 1.0 / scale;
  } else if (mode == DTWAX_NORM_MEDMAD) {
    auto kf = [&](int32_t i) -> uint64_t { return fkey(x[i]); };
    double med;
    if (n & 1) {
      med = (double)fkey_inv((uint32_t)radix_select<32>(n, n / 2, kf, hist, shv));
    } else {
      double a = (double)fkey_inv((uint32_t)radix_select<32>(n, n / 2 - 1, kf, hist, shv));
      double b = (double)fkey_inv((uint32_t)radix_select<32>(n, n / 2, kf, hist, shv));
      med = 0.5 * (a + b);
    }
    auto kd = [&](int32_t i) -> uint64_t { return dkey(fabs((double)x[i] - med)); };
    double mad;
    if (n & 1) {
      mad = dkey_inv(radix_select<64>(n, n / 2, kd, hist, shv));
    } else {
      double a = dkey_inv(radix_select<64>(n, n / 2 - 1, kd, hist, shv));
      double b = dkey_inv(radix_select<64>(n, n / 2, kd, hist, shv));
      mad = 0.5 * (a + b);
    }
    centre = med;
    scale = 1.4826 * mad;
    inv = scale > 0.0 ? 1.0 / scale : 0.0;
  }
}

__global__ __launch_bounds__(NT) void normalize_kernel(const float* __restrict__ raw,
                                                       const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ len,
                                                       int mode, double eps,
                                                       float* __restrict__ out,
                                                       double* __restrict__ stats) {
  __shared__ double sh[NT / 64 + 1];
  __shared__ uint32_t hist[256];
  __shared__ uint64_t shv[2];
  const int32_t r = blockIdx.x;
  const int32_t n = len[r];
  const float* x = raw + off[r];
  float* y = out + off[r];
  if (n <= 0) return;
  double centre, scale, inv;
  read_stats(x, n, mode, eps, sh, hist, shv, centre, scale, inv);
  if (mode == DTWAX_NORM_NONE) {
    for (int32_t i = threadIdx.x; i < n; i += NT) y[i] = x[i];
  } else {
    for (int32_t i = threadIdx.x; i < n; i += NT) y[i] = (float)(((double)x[i] - centre) * inv);
  }
  if (stats && threadIdx.x == 0) {
    stats[2 * (int64_t)r] = centre;
    stats[2 * (int64_t)r + 1] = scale;
  }
}

// Read Until chunks: chunk i (chunk_len samples at raw + i * chunk_len) of
// the read in slot[i].  A read's first chunk (first[i] != 0) fixes the read's
// statistics (stored as centre, inv in slot_stats[slot]); later chunks apply
// them, so the chunks of a read are normalised as one query would be with
// the first chunk's statistics.
__global__ __launch_bounds__(NT) void ru_normalize_kernel(const float* __restrict__ raw,
                                                          int32_t chunk_len,
                                                          const int32_t* __restrict__ slot,
                                                          const int32_t* __restrict__ first,
                                                          int mode, double eps,
                                                          double* __restrict__ slot_stats,
                                                          float* __restrict__ out) {
  __shared__ double sh[NT / 64 + 1];
  __shared__ uint32_t hist[256];
  __shared__ uint64_t shv[2];
  __shared__ double cs[2];
  const int32_t c = blockIdx.x;
  const float* x = raw + (int64_t)c * chunk_len;
  float* y = out + (int64_t)c * chunk_len;
  double* st = slot_stats + 2 * (int64_t)slot[c];
  if (first[c]) {
    double centre, scale, inv;
    read_stats(x, chunk_len, mode, eps, sh, hist, shv, centre, scale, inv);
    if (threadIdx.x == 0) {
      st[0] = centre;
      st[1] = inv;
      cs[0] = centre;
      cs[1] = inv;
    }
  } else if (threadIdx.x == 0) {
    cs[0] = st[0];
    cs[1] = st[1];
  }
  __syncthreads();
  const double centre = cs[0], inv = cs[1];
  for (int32_t i = threadIdx.x; i < chunk_len; i += NT) y[i] = (float)(((double)x[i] - centre) * inv);
}

hipError_t launch_normalize(const float* raw, const int64_t* off, const int32_t* len, int32_t n_reads,
                            int mode, double eps, float* out, double* stats, hipStream_t st) {
  if (n_reads <= 0) return hipSuccess;
  hipLaunchKernelGGL(normalize_kernel, dim3(n_reads), dim3(NT), 0, st, raw, off, len, mode, eps, out,
                     stats);
  return hipGetLastError();
}

hipError_t launch_ru_normalize(const float* raw, int32_t n_chunks, int32_t chunk_len, const int32_t* slot,
                               const int32_t* first, int mode, double eps, double* slot_stats,
                               float* out, hipStream_t st) {
  if (n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(ru_normalize_kernel, dim3(n_chunks), dim3(NT), 0, st, raw, chunk_len, slot, first,
                     mode, eps, slot_stats, out);
  return hipGetLastError();
}

}  // namespace dtwax
