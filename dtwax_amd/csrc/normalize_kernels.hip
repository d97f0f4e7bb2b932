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
//            absolute deviations), 8 bits per pass with an LDS histogram
//            and a parallel scan; the upper middle of an even count is
//            found by one more pass (the next key), not a second selection.
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

// Inclusive prefix sum over the workgroup of one value per thread (NT = 256).
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  uint32_t add = 0;
  for (int i = 0; i < w; i++) add += wsum[i];
  return v + add;
}

// Shared scratch of the order-statistic selection.
struct SelectShared {
  uint32_t hist[256];
  uint32_t wsum[NT / 64];
  uint64_t red[NT / 64];
  uint64_t v[3];
};

// k-th smallest (0-based) key over n elements, keys computed by `keyf(i)`:
// 8 bits per pass, an LDS histogram of the candidates' next digit, and a
// parallel scan that finds the digit holding rank k.  *rank_eq receives k's
// rank among the keys equal to the result and *n_eq their number.
template <int BITS, class KeyF>
__device__ uint64_t radix_select(int32_t n, int64_t k, KeyF keyf, SelectShared& S, int64_t* rank_eq,
                                 int64_t* n_eq) {
  static_assert(NT == 256, "one histogram bin per thread");
  uint64_t prefix = 0, mask = 0;
  int64_t eq = 0;
  for (int shift = BITS - 8; shift >= 0; shift -= 8) {
    S.hist[threadIdx.x] = 0;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += NT) {
      const uint64_t key = keyf(i);
      if ((key & mask) == prefix) atomicAdd(&S.hist[(key >> shift) & 0xff], 1u);
    }
    __syncthreads();
    const uint32_t c = S.hist[threadIdx.x];
    const uint32_t incl = block_incl_scan(c, S.wsum);
    const uint32_t excl = incl - c;
    if ((int64_t)excl <= k && k < (int64_t)incl) {  // exactly one bin holds rank k
      S.v[0] = prefix | ((uint64_t)threadIdx.x << shift);
      S.v[1] = (uint64_t)(k - (int64_t)excl);
      S.v[2] = c;
    }
    __syncthreads();
    prefix = S.v[0];
    k = (int64_t)S.v[1];
    eq = (int64_t)S.v[2];
    mask |= (uint64_t)0xff << shift;
    __syncthreads();
  }
  *rank_eq = k;
  *n_eq = eq;
  return prefix;
}

// Smallest key above `key` (one pass and a workgroup min).
template <class KeyF>
__device__ uint64_t next_key(int32_t n, uint64_t key, KeyF keyf, SelectShared& S) {
  uint64_t m = ~0ull;
  for (int32_t i = threadIdx.x; i < n; i += NT) {
    const uint64_t kk = keyf(i);
    if (kk > key && kk < m) m = kk;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint64_t t = __shfl_xor(m, o, 64);
    m = t < m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) S.red[threadIdx.x >> 6] = m;
  __syncthreads();
  uint64_t r = S.red[0];
  for (int i = 1; i < NT / 64; i++) r = S.red[i] < r ? S.red[i] : r;
  __syncthreads();
  return r;
}

// The middle order statistics of n keys: (k, k) for odd n, (k - 1, k) for
// even n, k = n / 2; the upper one of an even pair is the lower one again
// when that key repeats, else the next key above it.
template <int BITS, class KeyF>
__device__ void middle_keys(int32_t n, KeyF keyf, SelectShared& S, uint64_t& lo, uint64_t& hi) {
  int64_t rk, ne;
  if (n & 1) {
    lo = hi = radix_select<BITS>(n, n / 2, keyf, S, &rk, &ne);
    return;
  }
  lo = radix_select<BITS>(n, n / 2 - 1, keyf, S, &rk, &ne);
  hi = (rk + 1 < ne) ? lo : next_key(n, lo, keyf, S);
}

// Statistics of one read: centre, scale and the multiplier inv applied as
// y = (float)((x - centre) * inv).  NONE gives centre 0, inv 1 (a copy).
__device__ void read_stats(const float* __restrict__ x, int32_t n, int mode, double eps, double* sh,
                           SelectShared& S, double& centre, double& scale, double& inv) {
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
    uint64_t lo, hi;
    middle_keys<32>(n, kf, S, lo, hi);
    const double med = 0.5 * ((double)fkey_inv((uint32_t)lo) + (double)fkey_inv((uint32_t)hi));
    auto kd = [&](int32_t i) -> uint64_t { return dkey(fabs((double)x[i] - med)); };
    middle_keys<64>(n, kd, S, lo, hi);
    const double mad = 0.5 * (dkey_inv(lo) + dkey_inv(hi));
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
  __shared__ SelectShared S;
  const int32_t r = blockIdx.x;
  const int32_t n = len[r];
  const float* x = raw + off[r];
  float* y = out + off[r];
  if (n <= 0) return;
  double centre, scale, inv;
  read_stats(x, n, mode, eps, sh, S, centre, scale, inv);
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
  __shared__ SelectShared S;
  __shared__ double cs[2];
  const int32_t c = blockIdx.x;
  const float* x = raw + (int64_t)c * chunk_len;
  float* y = out + (int64_t)c * chunk_len;
  double* st = slot_stats + 2 * (int64_t)slot[c];
  if (first[c]) {
    double centre, scale, inv;
    read_stats(x, chunk_len, mode, eps, sh, S, centre, scale, inv);
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
