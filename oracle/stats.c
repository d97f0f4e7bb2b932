/*
 * stats.c -- normalisation statistics and the reference squiggle build.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_numpy_sum restates numpy's float64 add-reduction order (pairwise
 * summation, numpy/_core/src/umath/loops_utils.h.src pairwise_sum, blocks of
 * 128 with 8 accumulators, over 8192-element buffers summed in order) so that scripts/sdtw.py's np.mean values can be
 * reproduced bit for bit.  Pinned by tests/test_oracle_stats.py against
 * numpy itself on arrays of every length class, and by scripts/log.
 *
 * orc_python_znorm restates scripts/sdtw.py:86-99 (int_sig = 0).  Pinned by
 * scripts/log (query column, 4096 values, bit-exact).
 *
 * orc_norm_meanstd restates the arithmetic of normalizer.cu:71-114 (cuDNN
 * BatchNormalizationForwardTraining, SPATIAL, biased variance, eps 1e-4)
 * with FP64 statistics.  cuDNN's float reduction order is closed and no
 * reference vector pins it: low-bit parity with the reference's normaliser
 * is UNPINNED; the formula is pinned by closed forms (zero mean, unit
 * biased variance after normalisation) and by scripts/log with eps = 0.
 *
 * orc_norm_medmad: median/MAD z-score (the configs' normalisation; not in
 * the reference).  Pinned by numpy.median on random inputs.
 *
 * orc_squiggle_reference restates load_reference.hpp:33-108 (float
 * sequential sums, sqrt(E[x^2]-mu^2), (level-mu)/sd in float) with a direct
 * 2-bit k-mer index in place of the std::map scans; the arithmetic order is
 * unchanged.  Pinned by scripts/log's reference columns (FP64 python form
 * for the values, same k-mer indexing) and by an independent numpy float32
 * sequential-sum restatement in tests/test_oracle_stats.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static double pairwise(const double* a, int64_t n) {
  if (n < 8) {
    double res = 0.0;
    for (int64_t i = 0; i < n; i++) res += a[i];
    return res;
  } else if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  } else {
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise(a, n2) + pairwise(a + n2, n - n2);
  }
}

/* numpy feeds the reduction to the inner loop in buffers of 8192 elements
 * (NPY_BUFSIZE) and accumulates the per-buffer pairwise sums in order. */
double orc_numpy_sum(const double* a, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i += 8192) s += pairwise(a + i, (n - i) < 8192 ? (n - i) : 8192);
  return s;
}

void orc_python_znorm(const double* x, int64_t n, double* out) {
  double mean = orc_numpy_sum(x, n) / (double)n;
  double* d2 = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; i++) {
    double d = x[i] - mean;
    d2[i] = d * d;
  }
  double var = orc_numpy_sum(d2, n) / (double)n;
  double sd = pow(var, 0.5);
  for (int64_t i = 0; i < n; i++) out[i] = (x[i] - mean) / sd;
  free(d2);
}

void orc_norm_meanstd(const float* x, int64_t n, double eps, float* out,
                      double* mean_out, double* sd_out) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += (double)x[i];
  double mean = s / (double)n;
  double v = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double d = (double)x[i] - mean;
    v += d * d;
  }
  v /= (double)n;
  double inv = 1.0 / sqrt(v + eps);
  for (int64_t i = 0; i < n; i++) out[i] = (float)(((double)x[i] - mean) * inv);
  if (mean_out) *mean_out = mean;
  if (sd_out) *sd_out = sqrt(v + eps);
}

static int cmp_f(const void* a, const void* b) {
  float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}
static int cmp_d(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

void orc_norm_medmad(const float* x, int64_t n, float* out, double* med_out,
                     double* mad_out) {
  float* t = (float*)malloc(sizeof(float) * (size_t)n);
  memcpy(t, x, sizeof(float) * (size_t)n);
  qsort(t, (size_t)n, sizeof(float), cmp_f);
  double med = (n & 1) ? (double)t[n / 2]
                       : 0.5 * ((double)t[n / 2 - 1] + (double)t[n / 2]);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; i++) d[i] = fabs((double)x[i] - med);
  qsort(d, (size_t)n, sizeof(double), cmp_d);
  double mad = (n & 1) ? d[n / 2] : 0.5 * (d[n / 2 - 1] + d[n / 2]);
  double scale = 1.4826 * mad;
  double inv = scale > 0.0 ? 1.0 / scale : 0.0;
  for (int64_t i = 0; i < n; i++) out[i] = (float)(((double)x[i] - med) * inv);
  if (med_out) *med_out = med;
  if (mad_out) *mad_out = mad;
  free(t);
  free(d);
}

static int base2(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
  }
}

static int kmer_code(const char* p, int64_t avail, int k) {
  if (avail < k) return -1; /* substr shorter than k never matches */
  int code = 0;
  for (int i = 0; i < k; i++) {
    int b = base2(p[i]);
    if (b < 0) return -1;
    code = (code << 2) | b;
  }
  return code;
}

int64_t orc_squiggle_reference(const char* fwd, const char* rev, int64_t seq_len,
                               int k, const float* model_means, int64_t ref_len,
                               float* out_f32) {
  const char* target[2] = {fwd, rev};
  int64_t skipped = 0;
  memset(out_f32, 0, sizeof(float) * (size_t)(2 * ref_len));
  for (int j = 0; j < 2; j++) {
    float mean1 = 0.0f, stdev1 = 0.0f;
    for (int64_t i = 0; i < ref_len; i++) {
      int code = kmer_code(target[j] + i, seq_len - i, k);
      if (code < 0) continue;
      float lv = model_means[code];
      mean1 += lv;
      stdev1 = stdev1 + lv * lv;
    }
    mean1 = mean1 / (float)ref_len;
    stdev1 = stdev1 / (float)ref_len;
    stdev1 = sqrtf(stdev1 - (mean1 * mean1));
    for (int64_t i = 0; i < ref_len; i++) {
      int code = kmer_code(target[j] + i, seq_len - i, k);
      if (code < 0) { skipped++; continue; }
      float coeff = (model_means[code] - mean1) / stdev1;
      out_f32[j * ref_len + i] = coeff;
    }
  }
  return skipped;
}
