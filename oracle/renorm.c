/*
 * renorm.c -- the renormalised binary16 form (DTWAX_F16_RENORM) of this
 * build, restated for one read against one strand.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no such form: it is this build's answer to the binary16
 * drift of the reference's half2 arithmetic (DTW.cu:37-39), with the same
 * per-op cell t = rn(s - q), t2 = rn(t*t), D = rn(t2 + min(top, diag)).
 * The query's rows run in bands of 64*S rows (S rows per lane of a 64-lane
 * wave), the first band front-padded with zero-cost virtual rows so the
 * last row ends a band.  Every lane lowers its rows by the lane's drift Kb,
 * folded into the cost of its first row (every S-th row of the band):
 * t2 = rn(t*t - Kb), one rounding (a fused multiply-add), in place of
 * rn(t*t); virtual rows have t = 0, so that row costs -Kb.  A band after
 * the first also lowers its top row, the row the previous band left, by
 * that row's exact minimum m as it is staged: a[c] = rn(a[c] - m) for every
 * column (column -1 stays +inf).  Kb starts at rn(drift * S) and after each
 * band but the last becomes rn(Kb + rn(m * 2^-6)), m that band's last-row
 * minimum (the drift it left behind, over 64 lane hops).  The amounts are
 * added back in binary32, band by band: off += (64 * Kb + m_prev); the score
 * is (float)min(last row) + off, at the lowest column of the minimum.
 *
 * The split schedule (orc_sdtw_f16_renorm_split, the automatic choice for
 * reads over 1536 rows): the first 513..1536 rows in bands of 64 x 8 rows,
 * the rest in bands of 64 x 16 rows, the per-lane drift doubled at the
 * switch (twice the rows per lane); the first minimum subtractions come
 * after at most 512 rows, and the long bands start from an adapted drift.
 *
 * Parity pinned by: with drift 0 and one band it equals orc_sdtw_f16; an
 * independent per-cell restatement (numpy float16, the fused first-row cost
 * in exact rationals) on small inputs with several bands
 * (tests/test_oracle_renorm.py); the GPU kernels of DTWAX_F16_RENORM bit for
 * bit at every segment width (tests/test_gpu_parity.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

static float h2f(uint16_t h) { return orc_f16_to_f32(h); }
static uint16_t f2h(float f) { return orc_f32_to_f16(f); }
static uint16_t hmin(uint16_t a, uint16_t b) { return h2f(a) <= h2f(b) ? a : b; }

/* The schedule: the first Q - rest_rows rows in bands of 64*S rows
 * (front-padded), then rest_rows rows (a multiple of 64*S2) in bands of
 * 64*S2 rows, the per-lane drift doubling (S2 = 2S) at the switch.  With
 * rest_rows = 0 every band has 64*S rows. */
static void renorm_schedule(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, int S, int S2,
                            int64_t rest_rows, float drift, float* score, int64_t* argmin, float* last_row,
                            float* offset) {
  const uint16_t HINF = 0x7c00;
  const int64_t BR = 64 * (int64_t)S, BR2 = 64 * (int64_t)S2;
  const int64_t Q1 = Q - rest_rows;
  const int64_t nb1 = (Q1 + BR - 1) / BR, nb = nb1 + rest_rows / BR2;
  const int64_t P = nb1 * BR - Q1;
  uint16_t* a = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  uint16_t* b = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(N + 1));
  a[0] = b[0] = HINF; /* column -1 */
  for (int64_t c = 0; c < N; c++) a[c + 1] = 0; /* row -1: free start */
  uint16_t Kb = f2h(drift * (float)S), prev_min = 0;
  float off = 0.0f;
  for (int64_t band = 0; band < nb; band++) {
    const int second = band >= nb1;
    const int Sb = second ? S2 : S;
    const int64_t BRb = 64 * (int64_t)Sb;
    if (band == nb1 && band > 0) Kb = f2h(2.0f * h2f(Kb)); /* the drift of twice the rows per lane */
    const uint16_t low = band > 0 ? prev_min : 0;
    if (band > 0)
      for (int64_t c = 0; c < N; c++) a[c + 1] = f2h(h2f(a[c + 1]) - h2f(low));
    const uint16_t negK = (uint16_t)(Kb ^ 0x8000u);
    for (int64_t i = 0; i < BRb; i++) {
      const int64_t r = second ? Q1 + (band - nb1) * BR2 + i : band * BR - P + i;
      const float qr = r >= 0 ? h2f(q[r]) : 0.0f;
      const int lane_first = (i % Sb) == 0;
      for (int64_t c = 0; c < N; c++) {
        const uint16_t t = r >= 0 ? f2h(h2f(s[c]) - qr) : 0;
        const float tf = h2f(t);
        const uint16_t t2 = lane_first ? orc_h_fma(t, t, negK) : f2h(tf * tf);
        const uint16_t m = hmin(a[c + 1], a[c]);
        b[c + 1] = f2h(h2f(t2) + h2f(m));
      }
      uint16_t* tmp = a;
      a = b;
      b = tmp;
    }
    /* the amounts this band subtracted on every path, in binary32 as the
     * kernel adds them */
    const float t1 = 64.0f * h2f(Kb);
    const float t2 = t1 + h2f(low);
    off = off + t2;
    if (band + 1 < nb) {
      uint16_t m = HINF;
      for (int64_t c = 0; c < N; c++) m = hmin(m, a[c + 1]);
      prev_min = m;
      Kb = f2h(h2f(Kb) + h2f(f2h(h2f(m) * 0.015625f)));
    }
  }
  float best = INFINITY;
  int64_t bi = -1;
  for (int64_t c = 0; c < N; c++)
    if (h2f(a[c + 1]) < best) { best = h2f(a[c + 1]); bi = c; }
  *score = bi >= 0 ? best + off : INFINITY;
  *argmin = bi;
  if (last_row)
    for (int64_t c = 0; c < N; c++) last_row[c] = h2f(a[c + 1]);
  if (offset) *offset = off;
  free(a);
  free(b);
}

void orc_sdtw_f16_renorm(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, int S,
                         float drift, float* score, int64_t* argmin) {
  renorm_schedule(q, Q, s, N, S, S, 0, drift, score, argmin, NULL, NULL);
}

void orc_sdtw_f16_renorm_row(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, int S, int split,
                             float drift, float* last_row, float* offset) {
  float sc;
  int64_t am;
  renorm_schedule(q, Q, s, N, S, 2 * S, split ? orc_renorm_split_rows(Q) : 0, drift, &sc, &am, last_row,
                  offset);
}

int64_t orc_renorm_split_rows(int64_t Q) {
  /* rows in 1024-row bands after a first part of 513..1536 rows in 512-row
   * bands (reads of at most 1536 rows: all in 512-row bands) */
  return Q <= 1536 ? 0 : 1024 * ((Q - 513) / 1024);
}

void orc_sdtw_f16_renorm_split(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, float drift,
                               float* score, int64_t* argmin) {
  renorm_schedule(q, Q, s, N, 8, 16, orc_renorm_split_rows(Q), drift, score, argmin, NULL, NULL);
}

/* S = 0: the split schedule of orc_sdtw_f16_renorm_split */
void orc_batch_f16_renorm(const uint16_t* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                          const uint16_t* s, int64_t N, int n_strands, int S, float drift, float* score,
                          int64_t* argmin, int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_reads * n_strands; i++) {
    const int64_t r = i / n_strands, k = i % n_strands;
    if (S == 0)
      orc_sdtw_f16_renorm_split(q + offsets[r], lens[r], s + k * N, N, drift, &score[i], &argmin[i]);
    else
      orc_sdtw_f16_renorm(q + offsets[r], lens[r], s + k * N, N, S, drift, &score[i], &argmin[i]);
  }
}
