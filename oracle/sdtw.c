/*
 * sdtw.c -- clean sDTW recurrence in FP64 / FP32 / binary16, batched forms,
 * and the FP32 multi-core CPU baseline.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Recurrence (shipped build -DNO_REF_DEL, src/Makefile:8):
 *   D[r][c] = (s_c - q_r)^2 + min(D[r-1][c], D[r-1][c-1])
 * with row -1 = 0 (free start, DTW.cu:99-145 row-0 branch) and column -1 =
 * +inf (DTW.cu:241-244), score = min over the last row (DTW.cu:277-287,
 * 562-565).  Same recurrence as scripts/sdtw.py:205-225 in FP64.
 *
 * Per-op rounding of the reduced-precision forms follows COST_FUNCTION
 * (DTW.cu:37-39): FMA(FMA(SUB(r, q), SUB(r, q), 0), 1, MIN(t, d)), i.e.
 * t = rn(s - q), t2 = rn(t*t), D = rn(t2 + min(t, d)).
 *
 * Parity pinned by:
 *  - orc_sdtw_f64: scripts/log (every printed cell of the 4096 x 96,256 test
 *    matrix and the final min 286.8895631) -- tests/test_oracle_golden.py;
 *    full-matrix brute force on random tiny inputs; closed forms (Q = 1,
 *    exact substring => 0) -- tests/test_oracle_sdtw.py.
 *  - orc_sdtw_f32 / _f16: a per-cell restatement with numpy float32 /
 *    float16 scalar arithmetic on tiny inputs (every last-row cell), and
 *    agreement with FP64 to the precision's error bound.
 *  - orc_sdtw_f16_fused: a per-cell restatement in exact rationals (t*t + m
 *    formed exactly, one round-to-nearest-even to binary16 decided by exact
 *    comparison with the neighbouring binary16 values) on tiny inputs, every
 *    last-row cell; on 2 x 1 inputs where rounding t*t first changes the
 *    result, the fused and per-op forms differ and each equals its own
 *    restatement (tests/test_oracle_sdtw.py).
 *  - orc_cpu_baseline_f32: bit-equal to orc_sdtw_f32.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

/* ------------------------------------------------------------------ FP64 */
void orc_sdtw_f64(const double* q, int64_t Q, const double* s, int64_t N,
                  const double* init_row, double* last_row, double* score,
                  int64_t* argmin) {
  double* a = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double* b = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  for (int64_t c = 0; c < N; c++) a[c] = init_row ? init_row[c] : 0.0;
  for (int64_t r = 0; r < Q; r++) {
    const double qr = q[r];
    /* column 0: diag is column -1 = +inf, so min(top, diag) = top */
    if (N > 0) {
      double t = s[0] - qr;
      b[0] = t * t + a[0];
    }
    /* no dependency along the row (no reference deletion): vectorisable */
#pragma omp simd
    for (int64_t c = 1; c < N; c++) {
      double t = s[c] - qr;
      double top = a[c], left = a[c - 1];
      double m = top < left ? top : left;
      b[c] = t * t + m;
    }
    double* tmp = a; a = b; b = tmp;
  }
  double best = INFINITY;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (a[c] < best) { best = a[c]; bi = c; }
  if (Q == 0) { best = init_row ? best : 0.0; }
  if (last_row) memcpy(last_row, a, sizeof(double) * (size_t)N);
  *score = best;
  *argmin = bi;
  free(a);
  free(b);
}

void orc_sdtw_f64_trace(const double* q, int64_t Q, const double* s, int64_t N,
                        const int64_t* cols, int ncols, double* trace) {
  /* Same recurrence as orc_sdtw_f64; records D[r][cols[k]] for every row,
   * the cells scripts/sdtw.py:220-223 prints. */
  double* a = (double*)malloc(sizeof(double) * (size_t)N);
  double* b = (double*)malloc(sizeof(double) * (size_t)N);
  for (int64_t c = 0; c < N; c++) a[c] = 0.0;
  for (int64_t r = 0; r < Q; r++) {
    const double qr = q[r];
    double left = INFINITY;
    for (int64_t c = 0; c < N; c++) {
      double t = s[c] - qr;
      double top = a[c];
      double m = top < left ? top : left;
      b[c] = t * t + m;
      left = top;
    }
    for (int k = 0; k < ncols; k++) trace[r * ncols + k] = b[cols[k]];
    double* tmp = a; a = b; b = tmp;
  }
  free(a);
  free(b);
}

void orc_sdtw_bruteforce_f64(const double* q, int64_t Q, const double* s,
                             int64_t N, double* mat) {
  /* scripts/sdtw.py:205-219, literally. */
  for (int64_t i = 0; i < Q; i++)
    for (int64_t j = 0; j < N; j++) {
      double t = q[i] - s[j];
      double cost = pow(t, 2.0);
      if (i == 0) mat[j] = cost;
      else if (j == 0) mat[i * N] = mat[(i - 1) * N] + cost;
      else {
        double d = mat[(i - 1) * N + j - 1], u = mat[(i - 1) * N + j];
        mat[i * N + j] = cost + (d < u ? d : u);
      }
    }
}

/* ------------------------------------------------------------------ FP32 */
void orc_sdtw_f32(const float* q, int64_t Q, const float* s, int64_t N,
                  const float* init_row, float* last_row, float* score,
                  int64_t* argmin) {
  float* a = (float*)malloc(sizeof(float) * (size_t)(N + 1));
  float* b = (float*)malloc(sizeof(float) * (size_t)(N + 1));
  /* a[0] is column -1 (+inf); a[c+1] is column c */
  a[0] = b[0] = INFINITY;
  for (int64_t c = 0; c < N; c++) a[c + 1] = init_row ? init_row[c] : 0.0f;
  for (int64_t r = 0; r < Q; r++) {
    const float qr = q[r];
    for (int64_t c = 0; c < N; c++) {
      float t = s[c] - qr;
      float t2 = t * t;
      float top = a[c + 1], dg = a[c];
      float m = top < dg ? top : dg;
      b[c + 1] = t2 + m;
    }
    float* tmp = a; a = b; b = tmp;
  }
  float best = INFINITY;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (a[c + 1] < best) { best = a[c + 1]; bi = c; }
  if (last_row) memcpy(last_row, a + 1, sizeof(float) * (size_t)N);
  *score = best;
  *argmin = bi;
  free(a);
  free(b);
}

/* -------------------------------------------------------------- binary16 */
static float g_h2f[65536];
static int g_h2f_ready = 0;
static void h2f_init(void) {
  if (g_h2f_ready) return;
  for (int i = 0; i < 65536; i++) g_h2f[i] = orc_f16_to_f32((uint16_t)i);
  g_h2f_ready = 1;
}

static void sdtw_f16_impl(const uint16_t* q, int64_t Q, const uint16_t* s,
                          int64_t N, const uint16_t* init_row,
                          uint16_t* last_row, uint16_t* score, int64_t* argmin,
                          int fused) {
  h2f_init();
  uint16_t* a = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  uint16_t* b = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  const uint16_t HINF = 0x7c00;
  a[0] = b[0] = HINF;
  for (int64_t c = 0; c < N; c++) a[c + 1] = init_row ? init_row[c] : 0;
  for (int64_t r = 0; r < Q; r++) {
    const float qr = g_h2f[q[r]];
    for (int64_t c = 0; c < N; c++) {
      uint16_t t = orc_f32_to_f16(g_h2f[s[c]] - qr);
      uint16_t top = a[c + 1], dg = a[c];
      /* min of two binary16 values: compare as floats (exact) */
      uint16_t m = g_h2f[top] <= g_h2f[dg] ? top : dg;
      if (!fused) {
        float tf = g_h2f[t];
        uint16_t t2 = orc_f32_to_f16(tf * tf);
        b[c + 1] = orc_f32_to_f16(g_h2f[t2] + g_h2f[m]);
      } else {
        b[c + 1] = orc_h_fma(t, t, m);
      }
    }
    uint16_t* tmp = a; a = b; b = tmp;
  }
  float best = INFINITY;
  uint16_t bh = HINF;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (g_h2f[a[c + 1]] < best) { best = g_h2f[a[c + 1]]; bh = a[c + 1]; bi = c; }
  if (last_row) memcpy(last_row, a + 1, sizeof(uint16_t) * (size_t)N);
  *score = bh;
  *argmin = bi;
  free(a);
  free(b);
}

#if defined(__F16C__) && defined(__AVX2__)
#include <immintrin.h>
/* Same arithmetic as sdtw_f16_impl(fused = 0), eight columns at a time with
 * the x86 F16C conversions (IEEE round-to-nearest-even, subnormals kept).
 * Cross-checked against the scalar path by tests/test_oracle_sdtw.py. */
static void sdtw_f16_vec(const uint16_t* q, int64_t Q, const uint16_t* s,
                         int64_t N, const uint16_t* init_row, uint16_t* last_row,
                         uint16_t* score, int64_t* argmin) {
  h2f_init();
  const int64_t NP = (N + 7) / 8 * 8;
  /* a[0] = column -1 (+inf); a[c+1] = column c; padded to NP+8 */
  uint16_t* a = (uint16_t*)aligned_alloc(32, sizeof(uint16_t) * (size_t)(NP + 16));
  uint16_t* b = (uint16_t*)aligned_alloc(32, sizeof(uint16_t) * (size_t)(NP + 16));
  uint16_t* sp = (uint16_t*)aligned_alloc(32, sizeof(uint16_t) * (size_t)(NP + 16));
  const uint16_t HINF = 0x7c00;
  for (int64_t c = 0; c < NP + 16; c++) { a[c] = HINF; b[c] = HINF; sp[c] = 0; }
  for (int64_t c = 0; c < N; c++) { a[c + 1] = init_row ? init_row[c] : 0; sp[c] = s[c]; }
  const int R = _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC;
  for (int64_t r = 0; r < Q; r++) {
    const __m256 qv = _mm256_set1_ps(g_h2f[q[r]]);
    b[0] = HINF;
    for (int64_t c = 0; c < NP; c += 8) {
      __m256 sv = _mm256_cvtph_ps(_mm_loadu_si128((const __m128i*)(sp + c)));
      __m256 t = _mm256_cvtph_ps(_mm256_cvtps_ph(_mm256_sub_ps(sv, qv), R));
      __m256 t2 = _mm256_cvtph_ps(_mm256_cvtps_ph(_mm256_mul_ps(t, t), R));
      __m256 top = _mm256_cvtph_ps(_mm_loadu_si128((const __m128i*)(a + c + 1)));
      __m256 dg = _mm256_cvtph_ps(_mm_loadu_si128((const __m128i*)(a + c)));
      __m256 m = _mm256_min_ps(top, dg);
      __m128i v = _mm256_cvtps_ph(_mm256_add_ps(t2, m), R);
      _mm_storeu_si128((__m128i*)(b + c + 1), v);
    }
    uint16_t* tmp = a; a = b; b = tmp;
  }
  float best = INFINITY;
  uint16_t bh = HINF;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (g_h2f[a[c + 1]] < best) { best = g_h2f[a[c + 1]]; bh = a[c + 1]; bi = c; }
  if (last_row) memcpy(last_row, a + 1, sizeof(uint16_t) * (size_t)N);
  *score = bh;
  *argmin = bi;
  free(a);
  free(b);
  free(sp);
}
#endif

void orc_sdtw_f16(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N,
                  const uint16_t* init_row, uint16_t* last_row, uint16_t* score,
                  int64_t* argmin) {
#if defined(__F16C__) && defined(__AVX2__)
  sdtw_f16_vec(q, Q, s, N, init_row, last_row, score, argmin);
#else
  sdtw_f16_impl(q, Q, s, N, init_row, last_row, score, argmin, 0);
#endif
}

void orc_sdtw_f16_scalar(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N,
                         const uint16_t* init_row, uint16_t* last_row,
                         uint16_t* score, int64_t* argmin) {
  sdtw_f16_impl(q, Q, s, N, init_row, last_row, score, argmin, 0);
}
void orc_sdtw_f16_fused(const uint16_t* q, int64_t Q, const uint16_t* s,
                        int64_t N, const uint16_t* init_row, uint16_t* last_row,
                        uint16_t* score, int64_t* argmin) {
  sdtw_f16_impl(q, Q, s, N, init_row, last_row, score, argmin, 1);
}

/* --------------------------------------------------------------- batches */
static int set_threads(int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
  return omp_get_max_threads();
#else
  (void)n_threads;
  return 1;
#endif
}

void orc_batch_f64(const double* q, const int64_t* offsets, const int64_t* lens,
                   int64_t n_reads, const double* s, int64_t N, int n_strands,
                   double* score, int64_t* argmin, int n_threads) {
  set_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    int64_t r = i / n_strands, k = i % n_strands;
    orc_sdtw_f64(q + offsets[r], lens[r], s + k * N, N, NULL, NULL, &score[i],
                 &argmin[i]);
  }
}

void orc_batch_f32(const float* q, const int64_t* offsets, const int64_t* lens,
                   int64_t n_reads, const float* s, int64_t N, int n_strands,
                   float* score, int64_t* argmin, int n_threads) {
  set_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    int64_t r = i / n_strands, k = i % n_strands;
    orc_sdtw_f32(q + offsets[r], lens[r], s + k * N, N, NULL, NULL, &score[i],
                 &argmin[i]);
  }
}

void orc_batch_f16(const uint16_t* q, const int64_t* offsets,
                   const int64_t* lens, int64_t n_reads, const uint16_t* s,
                   int64_t N, int n_strands, int fused, uint16_t* score,
                   int64_t* argmin, int n_threads) {
  h2f_init();
  set_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    int64_t r = i / n_strands, k = i % n_strands;
    if (fused)
      sdtw_f16_impl(q + offsets[r], lens[r], s + k * N, N, NULL, NULL, &score[i],
                    &argmin[i], 1);
    else
      orc_sdtw_f16(q + offsets[r], lens[r], s + k * N, N, NULL, NULL, &score[i],
                   &argmin[i]);
  }
}

/* ------------------------------------------------- FP32 CPU baseline ----
 * Same arithmetic as orc_sdtw_f32 (built with -ffp-contract=off so t*t + m
 * is never fused), written so the column loop vectorises: the row update
 * has no intra-row dependency under NO_REF_DEL.  Long references are swept
 * in vertical strips of BASELINE_STRIP columns so the two rolling rows stay
 * in cache: a strip runs every query row, carrying its last column's value
 * per row (bnd[r] = D[r][c0 + W - 1]) into the next strip as that strip's
 * column c0 - 1 (the diagonal of its first column).  The cells and their
 * operations are the same as one pass over whole rows; strips are scanned
 * left to right, so the first minimum (lowest column) is kept.            */
#define BASELINE_STRIP 8192
static void baseline_one(const float* q, int64_t Q, const float* s, int64_t N,
                         float* a, float* b, float* bnd, float* score, int64_t* argmin) {
  const int64_t W = N > 2 * BASELINE_STRIP ? BASELINE_STRIP : N;
  float best = INFINITY;
  int64_t bi = -1;
  for (int64_t c0 = 0; c0 < N; c0 += W) {
    const int64_t w = (N - c0 < W) ? N - c0 : W;
    const float* sp = s + c0;
    /* row -1: 0 (free start); column c0 - 1 is +inf for the first strip */
    a[0] = c0 == 0 ? INFINITY : 0.0f;
    for (int64_t c = 0; c < w; c++) a[c + 1] = 0.0f;
    for (int64_t r = 0; r < Q; r++) {
      const float qr = q[r];
      const float* __restrict__ ap = a;
      float* __restrict__ bp = b + 1;
#pragma omp simd
      for (int64_t c = 0; c < w; c++) {
        float t = sp[c] - qr;
        float t2 = t * t;
        float top = ap[c + 1], dg = ap[c];
        float m = top < dg ? top : dg;
        bp[c] = t2 + m;
      }
      b[0] = c0 == 0 ? INFINITY : bnd[r]; /* D[r][c0 - 1]: next row's first diagonal */
      bnd[r] = b[w];                      /* D[r][c0 + w - 1] for the next strip */
      float* tmp = a; a = b; b = tmp;
    }
    for (int64_t c = 0; c < w; c++)
      if (a[c + 1] < best) { best = a[c + 1]; bi = c0 + c; }
  }
  *score = best;
  *argmin = bi;
}

int orc_cpu_baseline_f32(const float* q, const int64_t* offsets,
                         const int64_t* lens, int64_t n_reads, const float* s,
                         int64_t N, int n_strands, float* score,
                         int64_t* argmin, int n_threads) {
  int used = set_threads(n_threads);
  int64_t qmax = 1;
  for (int64_t r = 0; r < n_reads; r++) qmax = lens[r] > qmax ? lens[r] : qmax;
#pragma omp parallel
  {
    const int64_t W = N > 2 * BASELINE_STRIP ? BASELINE_STRIP : N;
    float* a = (float*)aligned_alloc(64, sizeof(float) * (size_t)((W + 1 + 15) / 16 * 16));
    float* b = (float*)aligned_alloc(64, sizeof(float) * (size_t)((W + 1 + 15) / 16 * 16));
    float* bnd = (float*)aligned_alloc(64, sizeof(float) * (size_t)((qmax + 15) / 16 * 16));
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < n_reads * n_strands; i++) {
      int64_t r = i / n_strands, k = i % n_strands;
      baseline_one(q + offsets[r], lens[r], s + k * N, N, a, b, bnd, &score[i],
                   &argmin[i]);
    }
    free(a);
    free(b);
    free(bnd);
  }
  return used;
}
