/*
 * oracle.h -- CPU restatement of the DTWax sDTW classifier path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dtwax_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it, and only as the checker (or, for
 * cpu_baseline, as the timed CPU port).  The product path fails loudly when
 * its HIP library is missing instead of falling back to anything here.
 *
 * Contents (each function cites the reference file:line it restates):
 *   binary16 conversion            orc_f32_to_f16 / orc_f16_to_f32
 *   clean sDTW recurrence          orc_sdtw_f64 / _f32 / _f16 / _f16_fused
 *   three-move recurrence          orc_sdtw3_f64 / _f32 / _f16  (sdtw3.c)
 *   full-matrix brute force        orc_sdtw_bruteforce_f64
 *   literal reference kernel       orc_emulate_reference   (src/include/DTW.cu)
 *   renormalised binary16 form     orc_sdtw_f16_renorm     (renorm.c; no reference counterpart)
 *   numpy-order statistics         orc_numpy_sum / orc_python_znorm
 *   query normalisation            orc_norm_meanstd / orc_norm_medmad
 *   reference squiggle build       orc_squiggle_reference  (load_reference.hpp)
 *   FP32 multi-core CPU baseline   orc_cpu_baseline_f32
 *
 * Parity status per function is recorded in the header of each .c file.
 */
#ifndef DTWAX_ORACLE_H
#define DTWAX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- binary16 ---------------------------------------------------------- */
uint16_t orc_f32_to_f16(float f);   /* IEEE-754 round-to-nearest-even */
float orc_f16_to_f32(uint16_t h);
void orc_f32_to_f16_array(const float* in, uint16_t* out, int64_t n);
void orc_f16_to_f32_array(const uint16_t* in, float* out, int64_t n);
uint16_t orc_h_add(uint16_t a, uint16_t b);
uint16_t orc_h_sub(uint16_t a, uint16_t b);
uint16_t orc_h_mul(uint16_t a, uint16_t b);
uint16_t orc_h_fma(uint16_t a, uint16_t b, uint16_t c); /* exact a*b+c, one rounding */

/* ---- clean sDTW (no reference deletion) --------------------------------
 *   D[r][c] = (s_c - q_r)^2 + min(D[r-1][c], D[r-1][c-1])
 *   D[-1][c] = init_row[c] (NULL: 0, free start), D[r][-1] = +inf
 *   score = min_c D[Q-1][c], argmin = lowest such c.
 * last_row (optional, N values) receives D[Q-1][*].
 * Restates src/include/DTW.cu:37-39 (COST_FUNCTION under NO_REF_DEL) and
 * scripts/sdtw.py:205-225 (sdtw_default).                                  */
void orc_sdtw_f64(const double* q, int64_t Q, const double* s, int64_t N,
                  const double* init_row, double* last_row,
                  double* score, int64_t* argmin);
/* FP32: t = rn(s-q); t2 = rn(t*t); D = rn(t2 + min) (DTW.cu:37-39 with the
 * FP32 macros of common.hpp:47-57). */
void orc_sdtw_f32(const float* q, int64_t Q, const float* s, int64_t N,
                  const float* init_row, float* last_row,
                  float* score, int64_t* argmin);
/* binary16, every op rounded to nearest even (common.hpp:32-45 half2 ops). */
void orc_sdtw_f16(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N,
                  const uint16_t* init_row, uint16_t* last_row,
                  uint16_t* score, int64_t* argmin);
/* Same as orc_sdtw_f16, one cell at a time without SIMD conversions (used to
 * cross-check the vectorised path). */
void orc_sdtw_f16_scalar(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N,
                         const uint16_t* init_row, uint16_t* last_row,
                         uint16_t* score, int64_t* argmin);
/* binary16 fused form: t = rn(s-q); D = rn(t*t + min) (one rounding). */
void orc_sdtw_f16_fused(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N,
                        const uint16_t* init_row, uint16_t* last_row,
                        uint16_t* score, int64_t* argmin);
/* FP64 recurrence recording D[r][cols[k]] for all rows (trace[Q*ncols]). */
void orc_sdtw_f64_trace(const double* q, int64_t Q, const double* s, int64_t N,
                        const int64_t* cols, int ncols, double* trace);
/* Full-matrix FP64 form of scripts/sdtw.py:205-225, for tiny inputs; fills
 * mat[Q*N] row-major. */
void orc_sdtw_bruteforce_f64(const double* q, int64_t Q, const double* s,
                             int64_t N, double* mat);

/* Batched forms (OpenMP over reads).  Queries are packed back to back;
 * offsets[i] is the first sample of read i, lens[i] its length.  Strands are
 * separate reference sequences of N columns each (s = n_strands*N values).
 * Output per read and strand: score[i*n_strands+k], argmin[...].           */
void orc_batch_f64(const double* q, const int64_t* offsets, const int64_t* lens,
                   int64_t n_reads, const double* s, int64_t N, int n_strands,
                   double* score, int64_t* argmin, int n_threads);
void orc_batch_f32(const float* q, const int64_t* offsets, const int64_t* lens,
                   int64_t n_reads, const float* s, int64_t N, int n_strands,
                   float* score, int64_t* argmin, int n_threads);
void orc_batch_f16(const uint16_t* q, const int64_t* offsets, const int64_t* lens,
                   int64_t n_reads, const uint16_t* s, int64_t N, int n_strands,
                   int fused, uint16_t* score, int64_t* argmin, int n_threads);

/* FP32 CPU baseline: same arithmetic as orc_batch_f32, vectorised over
 * columns, OpenMP over reads.  Returns the number of threads used. */
int orc_cpu_baseline_f32(const float* q, const int64_t* offsets,
                         const int64_t* lens, int64_t n_reads, const float* s,
                         int64_t N, int n_strands, float* score,
                         int64_t* argmin, int n_threads);

/* ---- three moves (reference built without -DNO_REF_DEL, DTW.cu:23-27) ---
 *   D[r][c] = (s_c - q_r)^2 + min(D[r][c-1], D[r-1][c], D[r-1][c-1])
 * free start, column -1 = +inf; same per-op rounding as the forms above.  */
void orc_sdtw3_f64(const double* q, int64_t Q, const double* s, int64_t N, double* score,
                   int64_t* argmin);
void orc_sdtw3_f32(const float* q, int64_t Q, const float* s, int64_t N, float* score,
                   int64_t* argmin);
void orc_sdtw3_f16(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, uint16_t* score,
                   int64_t* argmin);
void orc_batch3_f32(const float* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                    const float* s, int64_t N, int n_strands, float* score, int64_t* argmin,
                    int n_threads);
void orc_batch3_f16(const uint16_t* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                    const uint16_t* s, int64_t N, int n_strands, uint16_t* score, int64_t* argmin,
                    int n_threads);

/* ---- literal emulator of the reference kernel --------------------------
 * src/include/DTW.cu:151-572 (DTW) with compute_segment (:43-146), launched
 * as in DTW.hpp:287-292, reference tiled as in src/main.cu:92-109, queries
 * converted as in main.cu:165-178.  fp16 = 1: half2 (fwd, rev) lanes
 * (-DFP16); fp16 = 0: float, single concatenated sequence.  fix_q1 / fix_q2
 * switch off the two seam deviations (band-seam +inf diagonal; wrong shared
 * index in middle tiles) so the emulator can be checked against the clean
 * recurrence.                                                              */
typedef struct {
  int query_len;    /* QUERY_LEN   (common.hpp:65)          */
  int prefix_len;   /* PREFIX_LEN  (common.hpp:66)          */
  int segment;      /* SEGMENT_SIZE (common.hpp:64)         */
  int warp;         /* WARP_SIZE   (common.hpp:61), power of 2 */
  int ref_len;      /* REF_LEN     (common.hpp:71-76)       */
  int fp16;
  int fix_q1;
  int fix_q2;
} orc_emu_geom;
/* ref: fp16: 2*ref_len halves as float values (fwd[ref_len] then rev[ref_len]);
 *      fp32: ref_len floats.  queries: n_reads*query_len normalised floats.
 * out: fp16: n_reads*2 (fwd, rev) half values as float; fp32: n_reads floats.
 * last_row (optional): n_reads*ref_len entries (2 floats each for fp16) of
 * device_last_row as the kernel leaves it (tiled index order; entries it
 * never writes are +inf).
 * Returns 0, or -1 on an unsupported geometry.                             */
int orc_emulate_reference(const orc_emu_geom* g, const float* ref,
                          const float* queries, int64_t n_reads, float* out,
                          float* last_row, int n_threads);

/* ---- renormalised binary16 form (renorm.c) ------------------------------ */
/* DTWAX_F16_RENORM of this build for one read and one strand over whole
 * strands: bands of 64*S rows, a drift subtracted at every lane boundary
 * (and the previous band's row minimum at lane 0), the amounts added back
 * in binary32.  score: float; argmin: lowest column of the minimum. */
void orc_sdtw_f16_renorm(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, int S,
                         float drift, float* score, int64_t* argmin);
/* the automatic schedule for reads over 1536 rows: the first Q - split rows
 * in bands of 64 x 8 rows, the last split rows in bands of 64 x 16 rows */
int64_t orc_renorm_split_rows(int64_t Q);
void orc_sdtw_f16_renorm_split(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, float drift,
                               float* score, int64_t* argmin);
/* The last row of either schedule (split: the split schedule from S = 8),
 * lowered (binary16 values as float), and the binary32 offset the score
 * adds to its minimum. */
void orc_sdtw_f16_renorm_row(const uint16_t* q, int64_t Q, const uint16_t* s, int64_t N, int S, int split,
                             float drift, float* last_row, float* offset);
/* S = 0: the split schedule */
void orc_batch_f16_renorm(const uint16_t* q, const int64_t* offsets, const int64_t* lens, int64_t n_reads,
                          const uint16_t* s, int64_t N, int n_strands, int S, float drift, float* score,
                          int64_t* argmin, int n_threads);

/* ---- statistics and normalisation -------------------------------------- */
/* numpy's pairwise summation order for a float64 reduction (what np.sum /
 * np.mean use on a contiguous array). */
double orc_numpy_sum(const double* a, int64_t n);
/* scripts/sdtw.py:86-99 discrete_normalize (int_sig = 0): population std, no
 * epsilon, numpy summation order, pow(x, 0.5). */
void orc_python_znorm(const double* x, int64_t n, double* out);
/* Mean/std z-score with FP64 statistics: y = (x-mu)/sqrt(var + eps), biased
 * variance (normalizer.cu:86-100 semantics, eps 1e-4 there). */
void orc_norm_meanstd(const float* x, int64_t n, double eps, float* out,
                      double* mean, double* sd);
/* Median/MAD z-score: y = (x - med) / (1.4826 * MAD), med and MAD as the
 * mean of the two middle order statistics for even n. */
void orc_norm_medmad(const float* x, int64_t n, float* out, double* med,
                     double* mad);

/* ---- reference squiggle (load_reference.hpp:33-108) ----------------------
 * model_means[4096] indexed by 2-bit k-mer code (A=0,C=1,G=2,T=3, first base
 * most significant).  Per strand over the first ref_len k-mers: float
 * sequential sums of the level and its square; mean, sd = sqrt(E[x^2]-mu^2);
 * coeff = (level - mu)/sd in float.  Non-ACGT k-mers add nothing, count in
 * the denominator, and leave their slot at 0.  out_f32[2*ref_len]: fwd then
 * rev.  Returns the number of k-mers skipped.                               */
int64_t orc_squiggle_reference(const char* fwd, const char* rev, int64_t seq_len,
                               int k, const float* model_means, int64_t ref_len,
                               float* out_f32);

#ifdef __cplusplus
}
#endif
#endif
