/*
 * emulator.c -- literal CPU emulation of the reference sDTW kernel.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/include/DTW.cu:151-572 (__global__ DTW, shipped build
 * -DFP16 -DNO_REF_DEL) lane by lane: one block = one read, WARP lanes in
 * lock step, each lane holding SEGMENT reference registers and SEGMENT
 * penalty registers; __shfl_up_sync / __shfl_down_sync are array shifts;
 * the shared penalty_last_col and the global device_last_row are plain
 * buffers indexed exactly as the reference indexes them, including the two
 * seam deviations (a +inf diagonal at the first row of every band after the
 * first, DTW.cu:179-180, 238-244, 315, 442; and the middle-tile flush index
 * wave-(2*WARP-1) at DTW.cu:385-388 where tile 0 uses wave-WARP at :265-267).
 * The shfl_down by the literal 31 at DTW.cu:311/438 is emulated as a shift by
 * WARP-1 (identical at WARP = 32, the only width the reference builds).
 *
 * Arithmetic: half2 lanes (x = fwd, y = rev) with binary16 round-to-nearest
 * per op (common.hpp:32-45), or float for the FP32 build (common.hpp:47-57).
 *
 * Parity pinned by: with fix_q1 = fix_q2 = 1 the emulator must equal the
 * clean recurrence (orc_sdtw_f16 / orc_sdtw_f32) on the first
 * REF_BATCH*SEGMENT*WARP columns bit for bit, over many geometries and
 * random inputs; with the fixes off, differences may appear only at the seam
 * cells the code reading predicts (hand-derived on tiny geometries), and the
 * score can only rise.  tests/test_oracle_emulator.py.  No output of the
 * reference kernel for an input shipped in the reference exists, so the
 * emulator's fidelity to the real GPU rests on these checks.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define MAXW 64
#define MAXSEG 64

typedef struct { float v[2]; } val2;

typedef struct {
  const orc_emu_geom* g;
  int nc; /* 2 for half2, 1 for float */
} emu_ctx;

static inline float rnd(const emu_ctx* e, float x) {
  return e->g->fp16 ? orc_f16_to_f32(orc_f32_to_f16(x)) : x;
}
static inline val2 vset(float a) { val2 r; r.v[0] = a; r.v[1] = a; return r; }
static inline val2 vmin(val2 a, val2 b) {
  val2 r;
  for (int k = 0; k < 2; k++) r.v[k] = a.v[k] < b.v[k] ? a.v[k] : b.v[k];
  return r;
}
/* COST_FUNCTION under NO_REF_DEL (DTW.cu:37-39):
 * FMA(FMA(SUB(r1, q), SUB(r1, q), 0), 1, FIND_MIN(t, d)) */
static inline val2 cost_fn(const emu_ctx* e, val2 q, val2 r1, val2 t, val2 d) {
  val2 out = {{0.0f, 0.0f}};
  for (int k = 0; k < e->nc; k++) {
    float s = rnd(e, r1.v[k] - q.v[k]);
    float s2 = rnd(e, s * s);
    float m = t.v[k] < d.v[k] ? t.v[k] : d.v[k];
    out.v[k] = rnd(e, s2 + m);
  }
  if (e->nc == 1) out.v[1] = out.v[0];
  return out;
}

typedef struct {
  val2 ref_coeff1[MAXSEG];
  val2 penalty_here[MAXSEG];
  val2 penalty_temp[2];
  val2 penalty_left, penalty_diag, query_val, new_query_val;
  val2 min_segment, last_col;
} lane_t;

/* compute_segment, DTW.cu:43-146 */
static void compute_segment(const emu_ctx* e, int wave, int tid, lane_t* L,
                            int query_batch) {
  const int SEG = e->g->segment;
  const val2 zero = vset(0.0f);
  L->penalty_temp[0] = L->penalty_here[0];
  if ((tid != (wave - 1)) || query_batch) {
    L->penalty_here[0] = cost_fn(e, L->query_val, L->ref_coeff1[0],
                                 L->penalty_here[0], L->penalty_diag);
    int lim = (SEG % 2 == 0) ? SEG - 2 : SEG - 1;
    for (int i = 1; i < lim; i += 2) {
      L->penalty_temp[1] = L->penalty_here[i];
      L->penalty_here[i] = cost_fn(e, L->query_val, L->ref_coeff1[i],
                                   L->penalty_here[i], L->penalty_temp[0]);
      L->penalty_temp[0] = L->penalty_here[i + 1];
      L->penalty_here[i + 1] = cost_fn(e, L->query_val, L->ref_coeff1[i + 1],
                                       L->penalty_here[i + 1], L->penalty_temp[1]);
    }
    if (SEG > 1 && SEG % 2 == 0)
      L->penalty_here[SEG - 1] =
          cost_fn(e, L->query_val, L->ref_coeff1[SEG - 1],
                  L->penalty_here[SEG - 1], L->penalty_temp[0]);
  } else {
    L->penalty_here[0] =
        cost_fn(e, L->query_val, L->ref_coeff1[0], zero, L->penalty_diag);
    int lim = (SEG % 2 == 0) ? SEG - 2 : SEG - 1;
    for (int i = 1; i < lim; i += 2) {
      L->penalty_temp[1] = L->penalty_here[i];
      L->penalty_here[i] = cost_fn(e, L->query_val, L->ref_coeff1[i], zero, zero);
      L->penalty_temp[0] = L->penalty_here[i + 1];
      L->penalty_here[i + 1] =
          cost_fn(e, L->query_val, L->ref_coeff1[i + 1], zero, zero);
    }
    if (SEG > 1 && SEG % 2 == 0)
      L->penalty_here[SEG - 1] =
          cost_fn(e, L->query_val, L->ref_coeff1[SEG - 1], zero, zero);
  }
}

/* warp shuffles over one val2 field, all lanes at once */
#define SHFL_UP(W, lanes, field)                                       \
  do {                                                                 \
    val2 tmp_[MAXW];                                                   \
    for (int j_ = 0; j_ < (W); j_++)                                   \
      tmp_[j_] = (j_ >= 1) ? (lanes)[j_ - 1].field : (lanes)[j_].field; \
    for (int j_ = 0; j_ < (W); j_++) (lanes)[j_].field = tmp_[j_];     \
  } while (0)
#define SHFL_DOWN(W, lanes, field, delta)                               \
  do {                                                                  \
    val2 tmp_[MAXW];                                                    \
    for (int j_ = 0; j_ < (W); j_++)                                    \
      tmp_[j_] = (j_ + (delta) < (W)) ? (lanes)[j_ + (delta)].field     \
                                      : (lanes)[j_].field;              \
    for (int j_ = 0; j_ < (W); j_++) (lanes)[j_].field = tmp_[j_];      \
  } while (0)

typedef struct {
  const float* ref; /* untiled: comp k at ref[k*ref_len + col] */
  const val2* query; /* this block's QUERY_LEN values + padding */
  int64_t query_avail; /* readable entries of query[] */
} blk_in;

static val2 ref_at(const emu_ctx* e, const blk_in* b, int tile, int lane, int i) {
  /* tiled h[l*TILE + i*W + j] = untiled[l*TILE + j*SEG + i] (main.cu:92-109) */
  const orc_emu_geom* g = e->g;
  int64_t col = (int64_t)tile * g->segment * g->warp + (int64_t)lane * g->segment + i;
  val2 r;
  r.v[0] = b->ref[col];
  r.v[1] = (e->nc == 2) ? b->ref[(int64_t)g->ref_len + col] : r.v[0];
  return r;
}
static val2 query_at(const blk_in* b, int64_t idx) {
  if (idx < b->query_avail) return b->query[idx];
  return vset(0.0f); /* past the 32-entry pad: never consumed */
}

static void emulate_block(const emu_ctx* e, const blk_in* b, float* out, float* last_row) {
  const orc_emu_geom* g = e->g;
  const int W = g->warp, SEG = g->segment, PREFIX = g->prefix_len;
  const int TILE = SEG * W;
  const int REF_BATCH = g->ref_len / TILE;
  const int QUERY_BATCH = g->query_len / PREFIX;
  const int NUM_WAVES = PREFIX + (TILE - 1) / SEG;
  const int NW_BY_W = (NUM_WAVES / W) * W;
  const int RES = SEG - 1;
  const val2 INF = vset(INFINITY);
  const int plc_size = PREFIX + NUM_WAVES + 2 * W;
  val2* plc = (val2*)malloc(sizeof(val2) * (size_t)plc_size);
  val2* dlr = (val2*)malloc(sizeof(val2) * (size_t)g->ref_len);
  lane_t* L = (lane_t*)calloc((size_t)W, sizeof(lane_t));
  for (int i = 0; i < plc_size; i++) plc[i] = INF;
  for (int i = 0; i < g->ref_len; i++) dlr[i] = INF;

  for (int j = 0; j < W; j++) L[j].min_segment = INF; /* DTW.cu:168 */
  /* last_col_penalty_shuffled is uninitialised in the reference (:169); it
   * is always overwritten by lane W-1 before any flush reads it. */
  for (int j = 0; j < W; j++) L[j].last_col = INF;

  for (int qb = 0; qb < QUERY_BATCH; qb++) {
    const int last_band = (qb == QUERY_BATCH - 1);
    const int64_t qbase = (int64_t)qb * PREFIX;
    /* ---------------- tile 0: DTW.cu:179-300 */
    for (int j = 0; j < W; j++) {
      L[j].penalty_left = INF;
      L[j].penalty_diag = INF;
      for (int i = 0; i < SEG; i++) L[j].penalty_here[i] = vset(0.0f);
      L[j].query_val = INF;
      L[j].new_query_val = query_at(b, qbase + j);
    }
    L[0].query_val = L[0].new_query_val;
    SHFL_DOWN(W, L, new_query_val, 1);
    for (int j = 0; j < W; j++)
      for (int i = 0; i < SEG; i++) {
        L[j].ref_coeff1[i] = ref_at(e, b, 0, j, i);
        if (QUERY_BATCH > 1 && qb > 0) L[j].penalty_here[i] = dlr[j + i * W];
      }
    /* previous band's last-row value at this tile's last column, which the
     * fixed form needs as the next tile's first diagonal (it is overwritten
     * in dlr by the current band before the next tile starts) */
    val2 carry_lastcol = L[W - 1].penalty_here[SEG - 1];
    if (g->fix_q1 && qb > 0)
      for (int j = 1; j < W; j++) L[j].penalty_left = L[j - 1].penalty_here[SEG - 1];

    for (int wave = 1; wave <= NUM_WAVES; wave++) {
      if (last_band) SHFL_UP(W, L, min_segment);
      for (int j = 0; j < W; j++)
        if ((wave - PREFIX) <= j && j <= (wave - 1))
          compute_segment(e, wave, j, &L[j], qb);
      if ((wave & (W - 1)) == 0)
        for (int j = 0; j < W; j++) L[j].new_query_val = query_at(b, qbase + wave + j);
      SHFL_UP(W, L, query_val);
      for (int j = 0; j < W; j++) L[j].penalty_diag = L[j].penalty_left;
      {
        val2 tmp[MAXW];
        for (int j = 0; j < W; j++)
          tmp[j] = (j >= 1) ? L[j - 1].penalty_here[SEG - 1] : L[j].penalty_here[SEG - 1];
        for (int j = 0; j < W; j++) L[j].penalty_left = tmp[j];
      }
      L[0].query_val = L[0].new_query_val;
      L[0].penalty_left = INF;
      SHFL_DOWN(W, L, new_query_val, 1);
      if (REF_BATCH > 1) {
        SHFL_DOWN(W, L, last_col, 1);
        L[W - 1].last_col = L[W - 1].penalty_here[RES];
        if ((wave >= 2 * W - 1) && ((wave & (W - 1)) == W - 1)) {
          for (int j = 0; j < W; j++) plc[(wave - (2 * W - 1)) + j] = L[j].last_col;
        } else if (wave >= NW_BY_W) {
          plc[wave - W] = L[W - 1].penalty_here[RES];
        }
      }
      if (wave >= PREFIX && last_band)
        for (int j = 0; j < W; j++)
          for (int i = 0; i < SEG; i++)
            L[j].min_segment = vmin(L[j].min_segment, L[j].penalty_here[i]);
    }
    if (QUERY_BATCH > 1)
      for (int j = 0; j < W; j++)
        for (int i = 0; i < SEG; i++) dlr[j + i * W] = L[j].penalty_here[i];

    /* ---------------- tiles 1 .. REF_BATCH-1: DTW.cu:305-555 */
    for (int rb = 1; rb < REF_BATCH; rb++) {
      const int is_last = (rb == REF_BATCH - 1);
      const int64_t slot = is_last ? (int64_t)g->ref_len - TILE : (int64_t)rb * TILE;
      if (last_band) SHFL_DOWN(W, L, min_segment, W - 1);
      val2 prev_tile_last = carry_lastcol;
      for (int j = 0; j < W; j++) {
        L[j].penalty_diag = INF;
        L[j].penalty_left = INF;
        for (int i = 0; i < SEG; i++)
          L[j].penalty_here[i] = (qb == 0) ? vset(0.0f) : dlr[slot + j + i * W];
        L[j].query_val = INF;
        L[j].new_query_val = query_at(b, qbase + j);
      }
      L[0].query_val = L[0].new_query_val;
      L[0].penalty_left = plc[0];
      SHFL_DOWN(W, L, new_query_val, 1);
      for (int j = 0; j < W; j++)
        for (int i = 0; i < SEG; i++) L[j].ref_coeff1[i] = ref_at(e, b, rb, j, i);
      carry_lastcol = L[W - 1].penalty_here[SEG - 1];
      if (g->fix_q1 && qb > 0) {
        L[0].penalty_diag = prev_tile_last;
        for (int j = 1; j < W; j++) L[j].penalty_left = L[j - 1].penalty_here[SEG - 1];
      }

      for (int wave = 1; wave <= NUM_WAVES; wave++) {
        for (int j = 0; j < W; j++)
          if ((wave - PREFIX) <= j && j <= (wave - 1))
            compute_segment(e, wave, j, &L[j], qb);
        if ((wave & (W - 1)) == 0)
          for (int j = 0; j < W; j++) L[j].new_query_val = query_at(b, qbase + wave + j);
        SHFL_UP(W, L, query_val);
        L[0].query_val = L[0].new_query_val;
        if (!is_last) {
          SHFL_DOWN(W, L, last_col, 1);
          L[W - 1].last_col = L[W - 1].penalty_here[RES];
          if ((wave >= 2 * W - 1) && ((wave & (W - 1)) == W - 1)) {
            for (int j = 0; j < W; j++) plc[(wave - (2 * W - 1)) + j] = L[j].last_col;
          } else if (wave >= NW_BY_W) {
            int idx = g->fix_q2 ? (wave - W) : (wave - (2 * W - 1));
            plc[idx] = L[W - 1].penalty_here[RES];
          }
        }
        SHFL_DOWN(W, L, new_query_val, 1);
        for (int j = 0; j < W; j++) L[j].penalty_diag = L[j].penalty_left;
        {
          val2 tmp[MAXW];
          for (int j = 0; j < W; j++)
            tmp[j] = (j >= 1) ? L[j - 1].penalty_here[SEG - 1] : L[j].penalty_here[SEG - 1];
          for (int j = 0; j < W; j++) L[j].penalty_left = tmp[j];
        }
        L[0].penalty_left = plc[wave];
        if (wave >= PREFIX && last_band) {
          for (int j = 0; j < W; j++)
            for (int i = 0; i < SEG; i++)
              L[j].min_segment = vmin(L[j].min_segment, L[j].penalty_here[i]);
          if (wave != NUM_WAVES) SHFL_UP(W, L, min_segment);
        }
      }
      if (QUERY_BATCH > 1)
        for (int j = 0; j < W; j++)
          for (int i = 0; i < SEG; i++) dlr[slot + j + i * W] = L[j].penalty_here[i];
    }
  }
  out[0] = L[W - 1].min_segment.v[0];
  if (e->nc == 2) out[1] = L[W - 1].min_segment.v[1];
  /* device_last_row after the kernel (the last pass's rows; entries the
   * reference never writes keep their +inf initialisation here) */
  if (last_row)
    for (int i = 0; i < g->ref_len; i++)
      for (int k = 0; k < e->nc; k++) last_row[(int64_t)i * e->nc + k] = dlr[i].v[k];
  free(plc);
  free(dlr);
  free(L);
}

int orc_emulate_reference(const orc_emu_geom* g, const float* ref,
                          const float* queries, int64_t n_reads, float* out,
                          float* last_row, int n_threads) {
  if (g->warp < 1 || g->warp > MAXW || (g->warp & (g->warp - 1))) return -1;
  if (g->segment < 1 || g->segment > MAXSEG) return -1;
  if (g->prefix_len < 1 || g->query_len % g->prefix_len) return -1;
  if (g->ref_len < g->segment * g->warp) return -1;
  {
    /* the reference requires NUM_WAVES > 2*WARP (common.hpp:67-68, 92-93):
     * below that its flush indices go negative */
    int nw = g->prefix_len + (g->segment * g->warp - 1) / g->segment;
    if ((nw / g->warp) * g->warp < 2 * g->warp) return -1;
  }
  emu_ctx e = {g, g->fp16 ? 2 : 1};
  const int W = g->warp;
  /* main.cu:165-178: q -> {q, q} (half RN in the FP16 build), W zero pads */
  int64_t total = n_reads * (int64_t)g->query_len + W;
  val2* qv = (val2*)malloc(sizeof(val2) * (size_t)total);
  for (int64_t i = 0; i < total; i++) {
    float x = (i < n_reads * (int64_t)g->query_len) ? queries[i] : 0.0f;
    if (g->fp16) x = orc_f16_to_f32(orc_f32_to_f16(x));
    qv[i] = vset(x);
  }
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t r = 0; r < n_reads; r++) {
    blk_in b;
    b.ref = ref;
    b.query = qv + r * g->query_len;
    b.query_avail = total - r * g->query_len;
    emulate_block(&e, &b, out + r * e.nc,
                  last_row ? last_row + r * (int64_t)g->ref_len * e.nc : NULL);
  }
  free(qv);
  return 0;
}
