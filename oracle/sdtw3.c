/*
 * sdtw3.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The three-move sDTW recurrence of the reference built without -DNO_REF_DEL
 * (src/include/DTW.cu:23-27, COST_FUNCTION with FIND_MIN(l, FIND_MIN(t, d))):
 *   D[r][c] = (s_c - q_r)^2 + min(D[r][c-1], D[r-1][c], D[r-1][c-1])
 *   D[-1][c] = 0 (free start), D[r][-1] = +inf, score = min_c D[Q-1][c],
 *   argmin = lowest such c.
 * Per-op rounding as the two-move forms: binary32 / binary16 t = rn(s - q),
 * t2 = rn(t * t), D = rn(t2 + min) (the min is exact).
 *
 * Pinned by (tests/test_oracle_sdtw3.py): enumeration of every monotone
 * path (down, diagonal, right) on tiny inputs in pure Python; the two-move
 * score as an upper bound; Q = 1 equals min_c (s_c - q_0)^2; a query equal
 * to a reference substring scores 0; binary32/binary16 against a per-cell
 * numpy restatement on small inputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

void orc_sdtw3_f64(const double* q, int64_t Q, const double* s, int64_t N, double* score,
                   int64_t* argmin) {
  double* a = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* b = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  a[0] = b[0] = INFINITY; /* column -1 */
  for (int64_t c = 0; c < N; c++) a[c + 1] = 0.0;
  for (int64_t r = 0; r < Q; r++) {
    for (int64_t c = 0; c < N; c++) {
      double t = s[c] - q[r];
      double m = a[c + 1] < a[c] ? a[c + 1] : a[c];
      if (b[c] < m) m = b[c]; /* left: this row, previous column */
      b[c + 1] = t * t + m;
    }
    double* tmp = a; a = b; b = tmp;
  }
  double best = INFINITY;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (a[c + 1] < best) { best = a[c + 1]; bi = c; }
  *score = best;
  *argmin = bi;
  free(a);
  free(b);
}

void orc_sdtw3_f32(const float* q, int64_t Q, const float* s, int64_t N, float* score,
                   int64_t* argmin) {
  float* a = (float*)malloc(sizeof(float) * (size_t)(N + 1));
  float* b = (float*)malloc(sizeof(float) * (size_t)(N + 1));
  a[0] = b[0] = INFINITY;
  for (int64_t c = 0; c < N; c++) a[c + 1] = 0.0f;
  for (int64_t r = 0; r < Q; r++) {
    for (int64_t c = 0; c < N; c++) {
      float t = s[c] - q[r];
      float t2 = t * t;
      float m = a[c + 1] < a[c] ? a[c + 1] : a[c];
      if (b[c] < m) m = b[c];
      b[c + 1] = t2 + m;
    }
    float* tmp = a; a = b; b = tmp;
  }
  float best = INFINITY;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (a[c + 1] < best) { best = a[c + 1]; bi = c; }
  *score = best;
  *argmin = bi;
  free(a);
  free(b);
}

void orc_sdtw3_f16(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, uint16_t* score,
                   int64_t* argmin) {
  uint16_t* a = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  uint16_t* b = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  const uint16_t HINF = 0x7c00;
  a[0] = b[0] = HINF;
  for (int64_t c = 0; c < N; c++) a[c + 1] = 0;
  for (int64_t r = 0; r < Q; r++) {
    const float qr = orc_f16_to_f32(q[r]);
    for (int64_t c = 0; c < N; c++) {
      uint16_t t = orc_f32_to_f16(orc_f16_to_f32(s[c]) - qr);
      float tf = orc_f16_to_f32(t);
      uint16_t t2 = orc_f32_to_f16(tf * tf);
      uint16_t m = orc_f16_to_f32(a[c + 1]) <= orc_f16_to_f32(a[c]) ? a[c + 1] : a[c];
      if (orc_f16_to_f32(b[c]) < orc_f16_to_f32(m)) m = b[c];
      b[c + 1] = orc_f32_to_f16(orc_f16_to_f32(t2) + orc_f16_to_f32(m));
    }
    uint16_t* tmp = a; a = b; b = tmp;
  }
  float best = INFINITY;
  uint16_t bh = HINF;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (orc_f16_to_f32(a[c + 1]) < best) { best = orc_f16_to_f32(a[c + 1]); bh = a[c + 1]; bi = c; }
  *score = bh;
  *argmin = bi;
  free(a);
  free(b);
}

void orc_batch3_f32(const float* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                    const float* s, int64_t N, int n_strands, float* score, int64_t* argmin,
                    int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    const int64_t r = i / n_strands, k = i % n_strands;
    orc_sdtw3_f32(q + offsets[r], lens[r], s + k * N, N, &score[i], &argmin[i]);
  }
}

void orc_batch3_f16(const uint16_t* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                    const uint16_t* s, int64_t N, int n_strands, uint16_t* score, int64_t* argmin,
                    int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    const int64_t r = i / n_strands, k = i % n_strands;
    orc_sdtw3_f16(q + offsets[r], lens[r], s + k * N, N, &score[i], &argmin[i]);
  }
}
