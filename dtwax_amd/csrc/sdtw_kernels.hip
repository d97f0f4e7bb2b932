// sdtw_kernels.hip -- gfx950 subsequence-DTW kernels.
//
// Recurrence (reference build -DNO_REF_DEL, src/include/DTW.cu:37-39):
//   D[r][c] = (s_c - q_r)^2 + min(D[r-1][c], D[r-1][c-1])
//   D[-1][c] = top row (0 for a free start), D[r][-1] = +inf
//   score = min_c D[Q-1][c]
//
// There is no horizontal move, so a row depends only on the row above.  One
// wavefront owns one work item (read, reference column range).  The query is
// register resident: lane L holds rows [band0 + L*S, band0 + (L+1)*S) of the
// current band (64*S rows per band), and the wave sweeps reference columns
// with a one-step skew per lane (lane L is at column x + t - L at step t).
// Per step a lane updates its S cells of one column:
//   t = fma(s, a_i, -q_i); D_i = t*t + min(D_{i-1}(new), D_{i-1}(old))
// where a_i = 1 for real rows and 0 for the virtual rows that pad the first
// band (their cost is 0, so they carry the free-start row through).  Cells
// pass between lanes with one DPP wave_shr:1 per step (the lane above's last
// cell becomes this lane's top input; last step's top input is the diagonal).
// The last row of the band (lane 63, register S-1) is gathered over 64 steps
// in an LDS strip (every lane stores its last cell at slot k + 63 - lane at
// step k of a block, so lane 63's store is the last one at slots 0..63; a
// store rides the LDS pipe, where a DPP move would take a VALU issue slot)
// and then flushed: to the running min/argmin (final band) or to a row
// buffer (next band, or the Read Until state).
//
// Reference columns and the top-boundary row are streamed HBM -> LDS ring
// (256 slots + an 80-slot mirror so a lane reads the 64 consecutive slots of
// a block, and the next step's slots one step ahead, without wrapping) two
// blocks ahead with coalesced loads; each step reads its column from LDS with
// an immediate offset.  No MFMA: this is a min-plus recurrence.
//
// Compiled with -ffp-contract=off (the binary16 form must round t*t and the
// sum separately, as the reference does) and -fno-honor-nans (no NaN ever
// reaches the min, and without it every min is preceded by a canonicalise).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "kernels.h"

namespace dtwax {

typedef _Float16 half_t;
typedef half_t h2 __attribute__((ext_vector_type(2)));

#define DPP_WAVE_SHL1 0x130
#define DPP_WAVE_SHR1 0x138

template <class T>
__device__ __forceinline__ T dpp_shr1(T old, T src) {
  static_assert(sizeof(T) == 4, "32-bit DPP");
  int r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src),
                                      DPP_WAVE_SHR1, 0xf, 0xf, false);
  return __builtin_bit_cast(T, r);
}
template <class T>
__device__ __forceinline__ T dpp_shl1(T old, T src) {
  static_assert(sizeof(T) == 4, "32-bit DPP");
  int r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src),
                                      DPP_WAVE_SHL1, 0xf, 0xf, false);
  return __builtin_bit_cast(T, r);
}

__device__ __forceinline__ h2 hsplat(half_t v) { return (h2){v, v}; }
__device__ __forceinline__ h2 hmin2(h2 a, h2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ h2 hfma2(h2 a, h2 b, h2 c) { return __builtin_elementwise_fma(a, b, c); }
// Three-input packed min: one v_pk_minimum3_f16 on gfx950.  IEEE minimum
// (NaN-propagating, -0 < +0) equals min here: no NaN reaches it and every D
// is a sum of squares, so no -0.
__device__ __forceinline__ h2 hmin3(h2 a, h2 b, h2 c) {
  return __builtin_elementwise_minimum(__builtin_elementwise_minimum(a, b), c);
}

// Order-preserving map of a float to uint32 (for the 64-bit min keys).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t make_key(float v, int32_t col) {
  return ((uint64_t)ordered_bits(v) << 32) | (uint32_t)col;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// Issue priority by progress.  A SIMD's VALU issues from its oldest ready
// wave first, so of the two or three sDTW waves sharing a SIMD the oldest
// runs nearly alone and the youngest barely moves until the others are done
// (measured per item: 11.7 / 22.2 / 32.8 ms to finish three equal items
// started together).  At the end of a launch that youngest wave then runs
// its last item alone, below the SIMD's shared rate, about 1 ms per launch.
// The binary16 sweeps therefore lower their priority as they move through a
// band (3 in its first quarter down to 0 in its last): whichever wave is
// behind issues first, and the waves of a SIMD finish together (48.5 kb x
// 4096 x 2000: 43.9 -> 43.2 ms; fused 34.0 -> 32.9, renormalised 46.5 ->
// 44.8).  Launches with no more waves than SIMDs, and the Read Until
// engine's batches, which share SIMDs with each other, leave it off.  The
// binary32 sweep keeps
// the default order: its ops pair only across two waves, and the same rule
// lost 9% there on 4096 reads.
__device__ __forceinline__ void progress_prio(int32_t T, int32_t nsteps) {
  const int q = (int)(((int64_t)T * 4) / (nsteps > 0 ? nsteps : 1));
  if (q <= 0) __builtin_amdgcn_s_setprio(3);
  else if (q == 1) __builtin_amdgcn_s_setprio(2);
  else if (q == 2) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(0);
}

// Work fetch and result publication without lane-divergent branches: every
// lane issues the atomic, lane 0 with the real operand and the others with
// the identity, and the wave takes lane 0's result.  (A lane-0-only branch
// around the fetch inside the persistent loop lets the control-flow
// structurizer split the loop so that lanes 1..63 re-read the broadcast
// while lane 0 is masked off.)
__device__ __forceinline__ int fetch_item(int32_t* counter, int lane) {
  const int v = atomicAdd(counter, lane == 0 ? 1 : 0);
  return __builtin_amdgcn_readlane(v, 0);
}
__device__ __forceinline__ void publish_min(unsigned long long* addr, uint64_t key, int lane) {
  atomicMin(addr, lane == 0 ? (unsigned long long)key : ~0ull);
}

// FIX selects the fixed-geometry form of a kernel (the reference-layout
// entry points): item i is read i over the whole strand [0, n_cols) from
// x = 0, read i is q_stride samples from i * q_stride, binary16 queries are
// half2 words {q_fwd, q_rev}, and the final band's last row also goes to
// KParams::last_row in the reference's tiled layout.  The plan kernels
// (FIX = false) read the plan's items, offsets and lengths and compile
// without any of it.
template <bool FIX>
__device__ __forceinline__ WorkItem load_item(const KParams& p, int idx) {
  if constexpr (!FIX) {
    return p.items[idx];
  } else {
    WorkItem it{};
    it.read = idx;
    it.x = 0;
    it.c0 = 0;
    it.c1 = p.n_cols;
    return it;
  }
}
// First sample and length of a read.
template <bool FIX>
__device__ __forceinline__ int64_t read_offset(const KParams& p, int32_t read) {
  if constexpr (FIX) return (int64_t)read * p.q_stride;
  else return p.q_off[read];
}
template <bool FIX>
__device__ __forceinline__ int32_t read_length(const KParams& p, int32_t read) {
  if constexpr (FIX) return p.q_stride;
  else return p.q_len[read];
}
// Index of column col in the reference's tiled row layout (a tile of 832
// columns holds lane j's 26 columns j*26 + i at j + 32*i; the last tile sits
// at lr_last, src/include/DTW.cu:293, :419, :546).
__device__ __forceinline__ int64_t tiled_index(int32_t ntiles, int32_t last, int32_t col) {
  const int32_t t = col / TILE_W, w = col - t * TILE_W;
  const int32_t j = w / REF_SEG, i = w - j * REF_SEG;
  return (int64_t)(t < ntiles - 1 ? t * TILE_W : last) + j + REF_WARP * i;
}

constexpr int RING = 256;        // ring slots
constexpr int MIRROR = 80;       // slots [0, MIRROR) are also kept at RING + slot
constexpr int RING_ALLOC = RING + MIRROR;
// Last-row strip of a sweep, in LDS after the top-row ring: at step k of a
// 64-step block lane L stores its last cell at slot k + 63 - L (slots 0..126);
// lane 63 stores slot k at step k after every other lane's store there, so
// after the block slot j holds the last row at the block's j-th column.  One
// strip per column of a step (the two-column sweep has two).
constexpr int HSTRIP = 128;

// ------------------------------------------------------------------------
// binary16 forms (DTWAX_F16, DTWAX_F16_FUSED, DTWAX_F16_RENORM)
// ------------------------------------------------------------------------
template <int S, int MODE>
struct F16Band {
  h2 Qn[S];  // {-q, -q}
  h2 A[S];   // {1, 1} real row, {0, 0} virtual row
};

// One step: Do = column c-1 values, Dn = column c values.
//
// gfx950 needs a wait state between v_pk_fma_f16 and a dependent VALU op.
// The row costs t_i = fma(s, a_i, -q_i) are therefore issued one group of G
// rows ahead of the dependent chain (t*t, min, add), with scheduling barriers
// so the scheduler cannot pull a consumer back next to its fma.
template <int S, int MODE>
__device__ __forceinline__ void f16_cells(const F16Band<S, MODE>& B, h2 s, h2 tin, h2 dg,
                                          const h2 (&Do)[S], h2 (&Dn)[S]) {
  constexpr int G = 4;
  h2 tp = tin, dd = dg;
  h2 t[S];
#pragma unroll
  for (int i = 0; i < G && i < S; i++) t[i] = hfma2(s, B.A[i], B.Qn[i]);
#pragma unroll
  for (int g0 = 0; g0 < S; g0 += G) {
#pragma unroll
    for (int i = g0 + G; i < g0 + 2 * G && i < S; i++) t[i] = hfma2(s, B.A[i], B.Qn[i]);
#pragma unroll
    for (int i = g0; i < g0 + G && i < S; i++) {
      h2 m = hmin2(tp, dd);
      h2 v;
      if constexpr (MODE == DTWAX_F16_FUSED) {
        v = hfma2(t[i], t[i], m);
      } else {
        h2 c = t[i] * t[i];
        v = c + m;
      }
      Dn[i] = v;
      dd = Do[i];
      tp = v;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Per-op binary16 forms (DTWAX_F16, DTWAX_F16_RENORM), software-pipelined:
// every dependent pair of the chain min -> add -> min -> ... is separated by
// one independent instruction, in source order pinned by scheduling
// barriers, so no wait state is needed between them:
//   cell i:  m = min(tp, dd);  c[i+1] = t[i+1]^2;  v = c[i] + m;  t[i+PG] = fma(...)
// The fma slot of the last PG cells computes rows 0..PG-1 of the NEXT
// column (s_next) into tc, which the next step starts from.
constexpr int PG = 2;

//
// M3 (the reference built without -DNO_REF_DEL, DTW.cu:23-27): the left move
// D[r][c-1] = D_i(old) joins the min, which becomes one three-input min in
// the chain's min slot (still 4 ops per cell).
//
// DTWAX_F16_RENORM: row 0 (the lane's first) costs rn(t*t - K), one fused op
// in place of the multiply (negK = -K).
template <int S, int MODE, bool M3>
__device__ __forceinline__ void f16_cells_pipe(const F16Band<S, MODE>& B, h2 s, h2 s_next, h2 (&tc)[PG],
                                               h2 tin, h2 dg, const h2 (&Do)[S], h2 (&Dn)[S], h2 negK) {
  h2 t[S], c[S];
#pragma unroll
  for (int i = 0; i < PG; i++) t[i] = tc[i];
  c[0] = (MODE == DTWAX_F16_RENORM) ? hfma2(t[0], t[0], negK) : t[0] * t[0];
  h2 dd = dg;
  __builtin_amdgcn_sched_barrier(0);
  h2 tp = tin;
#pragma unroll
  for (int i = 0; i < S; i++) {
    h2 m;
    if constexpr (M3) m = hmin3(tp, dd, Do[i]);
    else m = hmin2(tp, dd);
    __builtin_amdgcn_sched_barrier(0);
    if (i + 1 < S) c[i + 1] = t[i + 1] * t[i + 1];
    __builtin_amdgcn_sched_barrier(0);
    const h2 v = c[i] + m;
    __builtin_amdgcn_sched_barrier(0);
    if (i + PG < S) t[i + PG] = hfma2(s, B.A[i + PG], B.Qn[i + PG]);
    else tc[i + PG - S] = hfma2(s_next, B.A[i + PG - S], B.Qn[i + PG - S]);
    __builtin_amdgcn_sched_barrier(0);
    Dn[i] = v;
    dd = Do[i];
    tp = v;
  }
}

struct SweepIO {
  const h2* ref;   // column 0 of the padded reference
  const h2* trow;  // top-boundary row (absolute columns) or null for a free start
  h2* orow;        // row output (absolute columns) or null
  int32_t x, c0, c1;
  bool final_band;
  bool top_diag;     // lane 0 takes trow[x-1] as the diagonal of its first cell (resumed
                     // top row; needs no virtual rows in front of the band)
  h2* lrow;          // FIX, final band: the last row also in the reference's tiled
                     // layout (this read's row of KParams::last_row), or null
  int32_t lr_ntiles, lr_last;
  h2 tlow;           // DTWAX_F16_RENORM: subtracted from the top row as it is staged
                     // (the previous band's, or the stored row's, exact minimum)
  bool prio;         // issue priority by progress through the band (progress_prio)
};

// Sweep one band across columns [x, c1) (+ ramp).  Accumulates the min/argmin
// of the band's last row over [c0, c1) into best/pos when final_band.
// DTWAX_F16_RENORM: K is the lane's drift, folded into its first row's cost
// (f16_cells_pipe); the top row is lowered by io.tlow as it is staged;
// rowmin receives the band's last-row minimum over [x, c1).
template <int S, int MODE, bool M3>
__device__ __forceinline__ void f16_sweep(const F16Band<S, MODE>& B, const SweepIO& io, h2 K,
                                          h2* __restrict__ ring_s, h2* __restrict__ ring_t,
                                          int lane, h2& best, int32_t& posf, int32_t& posr,
                                          h2& rowmin) {
  const h2 INF = hsplat((half_t)__builtin_inff());
  const h2 ZERO = hsplat((half_t)0.0f);
  const int32_t x = io.x;
  const int32_t nsteps = ((io.c1 - x + 63) + 63) & ~63;
  const h2* __restrict__ ref = io.ref;
  const h2* __restrict__ trow = io.trow;

  // Prologue: columns [x-64, x+128) into slots [0, 192) (+ mirror of [0, MIRROR)).
  {
    h2 a = ref[x - 64 + lane], b = ref[x + lane], c = ref[x + 64 + lane];
    ring_s[lane] = a;
    ring_s[RING + lane] = a;
    ring_s[64 + lane] = b;
    if (64 + lane < MIRROR) ring_s[RING + 64 + lane] = b;
    ring_s[128 + lane] = c;
    h2 ta = ZERO, tb = ZERO, tc = ZERO;
    if (trow) { ta = trow[x - 64 + lane]; tb = trow[x + lane]; tc = trow[x + 64 + lane]; }
    if (MODE == DTWAX_F16_RENORM) {  // the top row's lowering, as it is staged
      ta = ta - io.tlow;
      tb = tb - io.tlow;
      tc = tc - io.tlow;
    }
    ring_t[lane] = ta;
    ring_t[RING + lane] = ta;
    ring_t[64 + lane] = tb;
    if (64 + lane < MIRROR) ring_t[RING + 64 + lane] = tb;
    ring_t[128 + lane] = tc;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();

  h2 D0[S], D1[S];
#pragma unroll
  for (int i = 0; i < S; i++) D0[i] = INF;
  h2 tin = INF, dg = INF;
  h2* const hstrip = ring_t + RING_ALLOC;  // the last-row strip
  h2* const hput = hstrip + (63 - lane);
  if (lane == 0) tin = ring_t[64];  // top row at column x for lane 0
  if (lane == 0 && io.top_diag) dg = ring_t[63];
  const h2 negK = -K;

  constexpr bool PIPE = MODE != DTWAX_F16_FUSED;
  // pipelined forms: column of step 0 and its first PG row costs
  h2 s_cur = ring_s[(64 - lane) & (RING - 1)];
  h2 tf_cur = ring_t[((64 - lane) & (RING - 1)) + 1];  // top input of step 1 (read a step ahead)
  h2 tc[PG];
#pragma unroll
  for (int i = 0; i < PG; i++) tc[i] = hfma2(s_cur, B.A[i], B.Qn[i]);

  for (int32_t T = 0; T < nsteps; T += 64) {
    if (io.prio) progress_prio(T, nsteps);
    // Stream block T+128 .. T+191 (written to the ring after this block).
    h2 pf_s = ref[x + T + 128 + lane];
    h2 pf_t = trow ? trow[x + T + 128 + lane] : ZERO;
    if (MODE == DTWAX_F16_RENORM) pf_t = pf_t - io.tlow;
    const int base = (64 - lane + T) & (RING - 1);
#pragma unroll 1
    for (int g = 0; g < 64; g += 8) {
      const h2* rs = ring_s + base + g;
      const h2* rt = ring_t + base + g;
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        {
          h2 tf;
          if constexpr (PIPE) {
            // next step's column and the top input after it, read a step
            // ahead (slot base + 65 at most: inside the mirrored ring)
            const h2 s_next = rs[u + 1];
            const h2 tf_next = rt[u + 2];
            f16_cells_pipe<S, MODE, M3>(B, s_cur, s_next, tc, tin, dg, D0, D1, negK);
            s_cur = s_next;
            tf = tf_cur;
            tf_cur = tf_next;
          } else {
            h2 s = rs[u];
            f16_cells<S, MODE>(B, s, tin, dg, D0, D1);
            tf = rt[u + 1];
          }
          dg = tin;
          tin = dpp_shr1(tf, D1[S - 1]);
          hput[g + u] = D1[S - 1];
        }
        {
          h2 tf;
          if constexpr (PIPE) {
            const h2 s_next = rs[u + 2];
            const h2 tf_next = rt[u + 3];
            f16_cells_pipe<S, MODE, M3>(B, s_cur, s_next, tc, tin, dg, D1, D0, negK);
            s_cur = s_next;
            tf = tf_cur;
            tf_cur = tf_next;
          } else {
            h2 s = rs[u + 1];
            f16_cells<S, MODE>(B, s, tin, dg, D1, D0);
            tf = rt[u + 2];
          }
          dg = tin;
          tin = dpp_shr1(tf, D0[S - 1]);
          hput[g + u + 1] = D0[S - 1];
        }
      }
    }
    // strip slot j = last-row value of column x + T + j - 63
    const h2 hist = hstrip[lane];
    const int32_t col = x + T + lane - 63;
    const bool in_item = (col >= x) && (col < io.c1);
    if (io.final_band) {
      if (col >= io.c0 && col < io.c1) {
        if (hist.x < best.x) { best.x = hist.x; posf = col; }
        if (hist.y < best.y) { best.y = hist.y; posr = col; }
      }
    }
    if (MODE == DTWAX_F16_RENORM && !io.final_band && in_item) rowmin = hmin2(rowmin, hist);
    // band scratch rows keep the whole item; the final row only [c0, c1), so
    // the items of one read write disjoint columns of a stored row
    if (io.orow && in_item && (!io.final_band || col >= io.c0)) io.orow[col] = hist;
    // Ring refill for columns x + T + 128 + lane.
    __builtin_amdgcn_wave_barrier();
    const int slot = (T + 192 + lane) & (RING - 1);
    ring_s[slot] = pf_s;
    ring_t[slot] = pf_t;
    if (slot < MIRROR) { ring_s[RING + slot] = pf_s; ring_t[RING + slot] = pf_t; }
    __builtin_amdgcn_wave_barrier();
  }
}

// ---- binary16, two columns per step ----------------------------------------
// Lane L at step t computes columns cA = x + 2t - 2L and cB = cA + 1 of its
// S rows (a two-column lane skew).  The A column depends on the previous
// step's B column (its diagonal) and the B column on the A column one row
// up, so B is computed one row behind A: the pair (A_k, B_(k-1)) are two
// independent chains, and every dependent op of one has three ops of the
// other chain or of the look-ahead row costs between it and its producer.
// This removes the fused form's wait states (its chain min -> fma has no
// third op per cell to put between links): 21.5 -> 23.1 T cells/s at S = 32
// on the 48.5 kb x 4096 x 2000 batch, 20.9 -> 22.4 at S = 16; the per-op form
// runs at S = 32 as before (17.9) and 4.5% faster at S = 16.  Below S = 16
// the per-op and fused forms keep the one-column sweep (S = 8: 49.3 against
// 54.4 ms on the same batch, fused 39.7 against 41.8), as does the
// three-move recurrence.  The renormalised form, whose lane drift is folded
// into each lane's first-row cost (a fused op that the one-column step
// issues right before its first add), runs the two-column sweep from S = 8
// up (S = 8: 49.5 -> 48.6 ms; its split schedule 47.1 -> 46.6) and takes the
// band's last-row minimum in the history flush.  Every form's cells and
// their operations are the one-column sweep's, so the results are identical.
template <int S, int MODE, bool M3>
constexpr bool use_band2() {
  if (M3) return false;
  if (MODE == DTWAX_F16_RENORM) return S >= 8;
  return S >= 16 && (MODE == DTWAX_F16 || MODE == DTWAX_F16_FUSED);
}
constexpr int RING2 = 512;    // slots; column offset o = c - x lives in slot (o + 128) & 511
constexpr int MIRROR2 = 136;  // slots [0, MIRROR2) are also kept at RING2 + slot
constexpr int RING2_ALLOC = RING2 + MIRROR2;

// QX (the reference's middle-tile flush index, below): on lanes with xdiag
// set, the A column's rows k >= 2 take the next lane's B column row k - 2 of
// this same step as their diagonal (made one iteration earlier and moved by
// one DPP wave_shl:1) instead of their own previous column's row k - 1.
//
// DTWAX_F16_RENORM: the lane's first row (k = 0) costs rn(t*t - K), one fused
// op in place of the multiply (negK = -K), which lowers the lane's rows by
// its drift K at no extra op.
template <int S, int MODE, bool QX = false>
__device__ __forceinline__ void f16_cells2(const F16Band<S, MODE>& B, h2 sA, h2 sB, h2 tinA, h2 tinB,
                                           h2 diagA, h2 (&Bc)[S], h2& lastA, bool xdiag = false,
                                           h2 negK = (h2){(half_t)0.0f, (half_t)0.0f}) {
  constexpr bool FUSED = MODE == DTWAX_F16_FUSED;
  constexpr bool RN = MODE == DTWAX_F16_RENORM;
  // cA/cB: the pair's row costs t*t (per-op form) or t (fused form)
  const h2 t0 = hfma2(sA, B.A[0], B.Qn[0]);  // == rn(s - q) for a real row
  h2 cA = FUSED ? t0 : (RN ? hfma2(t0, t0, negK) : t0 * t0);
  h2 cB = hsplat((half_t)0.0f);
  h2 a1 = tinA, a2 = tinA, b2 = tinB;  // A_(k-1), A_(k-2), B_(k-2)
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k <= S; k++) {
    h2 mA, mB, nA, nB;
    auto row_costs = [&]() {
      if (k + 1 < S) nA = hfma2(sA, B.A[k + 1], B.Qn[k + 1]);
      if (k < S) nB = hfma2(sB, B.A[k], B.Qn[k]);
    };
    // The first mins take the top inputs that the lane move just delivered.
    // In the per-op forms from S = 25 up the row costs go first, so the move
    // has time to land (S = 32: 44.7 -> 44.0 ms on the 48.5 kb x 4096 x 2000
    // batch); the fused form and narrower bands run faster without it.
    constexpr bool COSTS_FIRST = !FUSED && S >= 25;
    if (COSTS_FIRST && k == 0) {
      row_costs();
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (QX) {
      if (k < S) {
        h2 d = k == 0 ? diagA : Bc[k - 1];
        if (k >= 2) {
          const h2 nb = dpp_shl1(Bc[k - 1], Bc[k - 2]);  // next lane's B_(k-2), this step
          d = xdiag ? nb : d;
        }
        mA = hmin2(a1, d);
      }
    } else {
      if (k < S) mA = hmin2(a1, k == 0 ? diagA : Bc[k - 1]);  // A_k: top A_(k-1), diagonal Bprev_(k-1)
    }
    if (k >= 1) mB = hmin2(b2, a2);                          // B_(k-1): top B_(k-2), diagonal A_(k-2)
    if (!(COSTS_FIRST && k == 0)) row_costs();
    __builtin_amdgcn_sched_barrier(0);
    h2 vA;
    if (k < S) vA = FUSED ? hfma2(cA, cA, mA) : cA + mA;
    if (k >= 1) {
      const h2 vB = FUSED ? hfma2(cB, cB, mB) : cB + mB;
      Bc[k - 1] = vB;  // after A_k read the previous column's row k-1
      b2 = vB;
    }
    if (k + 1 < S) cA = FUSED ? nA : nA * nA;
    if (k < S) cB = FUSED ? nB : ((RN && k == 0) ? hfma2(nB, nB, negK) : nB * nB);
    if (k < S) {
      a2 = a1;
      a1 = vA;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  lastA = a1;
}

// QK: the seam deviations of the reference kernel (src/include/DTW.cu), whose
// schedule runs 2048 query rows as two passes of 1024 rows over tiles of 832
// columns, one 32-lane warp with 26 columns per lane:
//   Q1  row 1024 (the second pass's first row) takes +inf as its diagonal at
//       columns 832k (k >= 1) and 832k + 26 (k >= 0): lane 0's diagonal is
//       reset at every tile start (DTW.cu:315, 442) and lane 1's first
//       diagonal is read before it is written (DTW.cu:179-180, 238-240);
//   Q2  at columns 832k (k >= 2), rows 962..992 of either pass take
//       D[r + 30][c - 1] as their diagonal: middle tiles flush the last
//       column to penalty_last_col[wave - 63] (DTW.cu:385-388) where tile 0
//       uses wave - 32 (DTW.cu:265-267).
// Here the 2048 rows are one band (64 lanes x 32 rows, x = 0), so the
// deviating cells are: lane 32 row 0 (Q1); lanes 30 and 62 rows 2..31 and
// lanes 31 and 63 row 0 (Q2), all at even (A) columns.  Lane l reaches A
// column 832k at step 416k + l, so the six deviating steps per tile are
// t = 30, 31, 32, 45, 62, 63 (mod 416), which fall in four 4-step groups;
// those groups run a slower step that applies them, the other 100 of every
// 104 groups run the plain step.  Q2's source D[r + 30][c - 1] for lane 30
// is lane 31's B column of the same step, made one iteration before lane 30
// needs it (f16_cells2's QX form); for lane 31 it is its own previous B
// column's row 30.
constexpr int QK_PERIOD = TILE_W / 2;  // steps per tile (two columns per step)

// One band over columns [x, c1) (+ ramp), two columns per step; band-final
// min/argmin per strand over [c0, c1) into best/posf/posr, the band's last
// row to io.orow.
// DTWAX_F16_RENORM: K is the lane's drift, folded into its first row's cost
// (f16_cells2); the band's top row is lowered by io.tlow as it is staged;
// rowmin receives the band's last-row minimum over [x, c1) (not the final
// band).
template <int S, int MODE, bool QK = false, bool FIX = false>
__device__ __forceinline__ void f16_band2(const F16Band<S, MODE>& B, const SweepIO& io, h2* ring_s,
                                          h2* ring_t, int lane, h2& best, int32_t& posf, int32_t& posr,
                                          h2 K = (h2){(half_t)0.0f, (half_t)0.0f}, h2* rowmin = nullptr) {
  constexpr bool RN = MODE == DTWAX_F16_RENORM;
  static_assert(!QK || (S == 32 && MODE == DTWAX_F16 && FIX),
                "seam deviations: binary16 per op, 64 x 32 rows, whole strands from column 0");
  const h2 INF = hsplat((half_t)__builtin_inff());
  const h2 ZERO = hsplat((half_t)0.0f);
  const int32_t x = io.x, c0 = io.c0, c1 = io.c1;
  const h2* __restrict__ ref = io.ref;
  // top rows from trow, or 0 (free start): loads without branches
  const h2* tsrc = io.trow ? io.trow : ref;
  const h2 tmask = io.trow ? hsplat((half_t)1.0f) : ZERO;
  auto put = [&](h2* ring, int slot, h2 v) {
    ring[slot] = v;
    if (slot < MIRROR2) ring[RING2 + slot] = v;
  };
  // Prologue: offsets [-128, 0) are before the item (their cells are +inf
  // whatever the column holds): zeros; offsets [0, 130) from memory.
  put(ring_s, lane, ZERO);
  put(ring_s, 64 + lane, ZERO);
  put(ring_t, lane, ZERO);
  put(ring_t, 64 + lane, ZERO);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int o = 64 * k + lane;
    if (o < 130) {
      put(ring_s, o + 128, ref[x + o]);
      h2 tv = tsrc[x + o] * tmask;
      if constexpr (RN) tv = tv - io.tlow;
      put(ring_t, o + 128, tv);
    }
  }
  // a resumed row also gives lane 0 its first diagonal, trow[x - 1] (slot 127)
  if (io.top_diag && lane == 0) put(ring_t, 127, RN ? io.trow[x - 1] - io.tlow : io.trow[x - 1]);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();

  h2 Bc[S];
#pragma unroll
  for (int i = 0; i < S; i++) Bc[i] = INF;
  h2 tinA = INF, tinB = INF, diagA = INF;
  // the last-row strips of the A and B columns
  h2* const hstripA = ring_t + RING2_ALLOC;
  h2* const hstripB = hstripA + HSTRIP;
  h2* const hputA = hstripA + (63 - lane);
  h2* const hputB = hstripB + (63 - lane);
  if (lane == 0) {
    tinA = ring_t[128];
    tinB = ring_t[129];
    if (io.top_diag) diagA = ring_t[127];
  }
  const h2 negK = -K;
  h2 rmin = INF;
  const int base0 = (128 - 2 * lane) & (RING2 - 1);
  h2 sA = ring_s[base0], sB = ring_s[base0 + 1];

  const int32_t nsteps = (((c1 - x + 127) >> 1) + 63) & ~63;
  int32_t tmod = 0;  // QK: T mod QK_PERIOD
  for (int32_t T = 0; T < nsteps; T += 64) {
    if (io.prio) progress_prio(T, nsteps);
    // offsets [2T + 130, 2T + 258): written to the ring after this block
    h2 pf_s[2], pf_t[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int32_t c = x + 2 * T + 130 + 64 * h + lane;
      pf_s[h] = ref[c];
      pf_t[h] = tsrc[c] * tmask;
      if constexpr (RN) pf_t[h] = pf_t[h] - io.tlow;
    }
    const int base = (2 * T + 128 - 2 * lane) & (RING2 - 1);
#pragma unroll 1
    for (int g = 0; g < 64; g += 4) {
      const h2* rs = ring_s + base + 2 * g;
      const h2* rt = ring_t + base + 2 * g;
      if constexpr (QK) {
        int32_t r0 = tmod + g;
        if (r0 >= QK_PERIOD) r0 -= QK_PERIOD;
        if (r0 == 28 || r0 == 32 || r0 == 44 || r0 == 60) {
          // a group holding a deviating step (uniform across the wave).  Only
          // steps 30 and 62 (u = 2 of groups 28 and 60) need the cross-lane
          // diagonal (f16_cells2's QX form, S - 2 more lane moves); the
          // others change one lane's first diagonal, a select
          const bool xgroup = r0 == 28 || r0 == 60;
          auto step = [&](auto uc, auto qx) {
            constexpr int u = decltype(uc)::value;
            const int32_t t = T + g + u, r = r0 + u;
            const bool q1 = (r == 32 && t >= QK_PERIOD + 32) || r == 45;      // lane 32
            const bool q2 = (r == 30 || r == 31 || r == 62 || r == 63) && t >= 2 * QK_PERIOD + r;
            h2 dA = diagA;
            if (q1 && lane == 32) dA = INF;
            if (q2 && (r & 1) && lane == r) dA = Bc[S - 2];                  // lanes 31, 63: own row 30
            const h2 nA = rs[2 * u + 2], nB = rs[2 * u + 3];
            const h2 tfA = rt[2 * u + 2], tfB = rt[2 * u + 3];
            h2 lastA;
            if constexpr (decltype(qx)::value)
              f16_cells2<S, MODE, true>(B, sA, sB, tinA, tinB, dA, Bc, lastA, xgroup && q2 && lane == r);  // lanes 30, 62
            else
              f16_cells2<S, MODE>(B, sA, sB, tinA, tinB, dA, Bc, lastA);
            diagA = tinB;
            tinA = dpp_shr1(tfA, lastA);
            tinB = dpp_shr1(tfB, Bc[S - 1]);
            hputA[g + u] = lastA;
            hputB[g + u] = Bc[S - 1];
            sA = nA;
            sB = nB;
          };
          using F = std::false_type;
          step(std::integral_constant<int, 0>{}, F{});
          step(std::integral_constant<int, 1>{}, F{});
          step(std::integral_constant<int, 2>{}, std::true_type{});
          step(std::integral_constant<int, 3>{}, F{});
          continue;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        // next step's columns; the top inputs of the next step (lane 0)
        const h2 nA = rs[2 * u + 2], nB = rs[2 * u + 3];
        const h2 tfA = rt[2 * u + 2], tfB = rt[2 * u + 3];
        h2 lastA;
        f16_cells2<S, MODE>(B, sA, sB, tinA, tinB, diagA, Bc, lastA, false, negK);
        diagA = tinB;
        tinA = dpp_shr1(tfA, lastA);
        tinB = dpp_shr1(tfB, Bc[S - 1]);
        hputA[g + u] = lastA;
        hputB[g + u] = Bc[S - 1];
        sA = nA;
        sB = nB;
      }
    }
    // strip slot j holds the last row at columns x + 2(T + j) - 126 (A) and + 1 (B)
    const h2 histA = hstripA[lane], histB = hstripB[lane];
    const int32_t colA = x + 2 * (T + lane) - 126;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int32_t col = colA + h;
      const h2 v = h ? histB : histA;
      if (io.final_band && col >= c0 && col < c1) {
        if (v.x < best.x) { best.x = v.x; posf = col; }
        if (v.y < best.y) { best.y = v.y; posr = col; }
      }
      if (io.orow && col >= x && col < c1 && (!io.final_band || col >= c0)) io.orow[col] = v;
      if constexpr (RN)
        if (!io.final_band && col >= x && col < c1) rmin = hmin2(rmin, v);
      if constexpr (FIX)
        if (io.lrow && io.final_band && col >= c0 && col < c1)
          io.lrow[tiled_index(io.lr_ntiles, io.lr_last, col)] = v;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int slot = (2 * T + 258 + 64 * h + lane) & (RING2 - 1);
      put(ring_s, slot, pf_s[h]);
      put(ring_t, slot, pf_t[h]);
    }
    __builtin_amdgcn_wave_barrier();
    if constexpr (QK) {
      tmod += 64;
      if (tmod >= QK_PERIOD) tmod -= QK_PERIOD;
    }
  }
  if constexpr (RN)
    if (rowmin) *rowmin = rmin;
}

// The shared band of a pack: f16_band2 with every read's lanes independent.
// A read's first lane takes its top inputs from ring_t at its own columns
// (0 from the item's first column on, +inf before it: the free start that
// lane 0 of an unpacked band has), and its last lane inserts its last row
// into the read's history chain (lane 63 alone does so in an unpacked band).
// The chains are flushed every F steps, F at most the shortest read's lane
// count, so no value leaves its read's lanes unseen: PL_FINAL lanes fold the
// values into best/pos, PL_HAND lanes write them to the read's scratch row
// `hrow` for its full bands.  Lane L of a read ending at lane E holds, after
// step t, the value E made at step t - (E - L): column x + 2t - 4E + 2L (A).
template <int S, int MODE>
__device__ __forceinline__ void f16_band2_pack(const F16Band<S, MODE>& B, const SweepIO& io, h2* ring_s,
                                               h2* ring_t, int lane, int32_t info, h2* hrow, int32_t F,
                                               h2& best, int32_t& posf, int32_t& posr) {
  const h2 INF = hsplat((half_t)__builtin_inff());
  const h2 ZERO = hsplat((half_t)0.0f);
  const int32_t x = io.x, c0 = io.c0, c1 = io.c1;
  const h2* __restrict__ ref = io.ref;
  const bool start = (info & PL_START) != 0;
  const bool end = (info & 255) == lane;
  const bool fin = (info & PL_FINAL) != 0;
  const bool hand = (info & PL_HAND) != 0;
  const int32_t cst = x + 2 * lane - 4 * (info & 255);
  auto put = [&](h2* ring, int slot, h2 v) {
    ring[slot] = v;
    if (slot < MIRROR2) ring[RING2 + slot] = v;
  };
  put(ring_s, lane, ZERO);
  put(ring_s, 64 + lane, ZERO);
  put(ring_t, lane, INF);
  put(ring_t, 64 + lane, INF);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int o = 64 * k + lane;
    if (o < 130) {
      put(ring_s, o + 128, ref[x + o]);
      put(ring_t, o + 128, ZERO);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();

  h2 Bc[S];
#pragma unroll
  for (int i = 0; i < S; i++) Bc[i] = INF;
  h2 tinA = INF, tinB = INF, diagA = INF, histA = INF, histB = INF;
  if (lane == 0) {
    tinA = ring_t[128];
    tinB = ring_t[129];
  }
  const int base0 = (128 - 2 * lane) & (RING2 - 1);
  h2 sA = ring_s[base0], sB = ring_s[base0 + 1];

  auto flush = [&](int32_t t) {
    const int32_t colA = cst + 2 * t;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int32_t col = colA + h;
      const h2 v = h ? histB : histA;
      if (fin && col >= c0 && col < c1) {
        if (v.x < best.x) { best.x = v.x; posf = col; }
        if (v.y < best.y) { best.y = v.y; posr = col; }
      }
      if (hand && col >= x && col < c1) hrow[col] = v;
    }
  };

  const int32_t nsteps = (((c1 - x + 127) >> 1) + 63) & ~63;
  for (int32_t T = 0; T < nsteps; T += 64) {
    if (io.prio) progress_prio(T, nsteps);
    h2 pf_s[2];
#pragma unroll
    for (int h = 0; h < 2; h++) pf_s[h] = ref[x + 2 * T + 130 + 64 * h + lane];
    const int base = (2 * T + 128 - 2 * lane) & (RING2 - 1);
#pragma unroll 1
    for (int g = 0; g < 64; g += 4) {
      const h2* rs = ring_s + base + 2 * g;
      const h2* rt = ring_t + base + 2 * g;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const h2 nA = rs[2 * u + 2], nB = rs[2 * u + 3];
        const h2 tfA = rt[2 * u + 2], tfB = rt[2 * u + 3];
        h2 lastA;
        f16_cells2<S, MODE>(B, sA, sB, tinA, tinB, diagA, Bc, lastA);
        diagA = tinB;
        const h2 shA = dpp_shr1(tfA, lastA), shB = dpp_shr1(tfB, Bc[S - 1]);
        tinA = start ? tfA : shA;
        tinB = start ? tfB : shB;
        const h2 hA = dpp_shl1(lastA, histA), hB = dpp_shl1(Bc[S - 1], histB);
        histA = end ? lastA : hA;
        histB = end ? Bc[S - 1] : hB;
        sA = nA;
        sB = nB;
      }
      if (((g + 4) & (F - 1)) == 0) flush(T + g + 3);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int slot = (2 * T + 258 + 64 * h + lane) & (RING2 - 1);
      put(ring_s, slot, pf_s[h]);
      put(ring_t, slot, ZERO);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__device__ __forceinline__ half_t clamp_half(float v) { return (half_t)v; }

// The renormalised form's split schedule for reads over 1536 rows: the rows
// after the first 513..1536 (whose bands are 64 x 8 rows) run in bands of
// 64 x 16 rows, so the first exact minimum subtraction comes within 512
// rows and the long bands start from an adapted drift.
__device__ __forceinline__ int32_t renorm_split_rows(int32_t Q) {
  return Q <= 1536 ? 0 : 1024 * ((Q - 513) / 1024);
}

// One band of SB rows per lane, its first lane's first row being read row r0
// (rows < 0 are virtual: zero cost).
template <int SB, int MODE, bool M3, bool QK, bool FIX>
__device__ __forceinline__ void f16_item_band(const float* __restrict__ qp, const uint32_t* __restrict__ qw,
                                              int32_t r0, const SweepIO& io, h2 Kl, h2* ring_s, h2* ring_t,
                                              int lane, h2& best, int32_t& posf, int32_t& posr, h2& rowmin) {
  F16Band<SB, MODE> B;
#pragma unroll
  for (int i = 0; i < SB; i++) {
    const int32_t r = r0 + i;
    if (r >= 0) {
      // queries are read once: a non-temporal load keeps them from evicting
      // the reference, which every work item re-reads, from L2
      if constexpr (FIX) {
        B.Qn[i] = -__builtin_bit_cast(h2, __builtin_nontemporal_load(&qw[r]));
      } else {
        half_t q = (half_t)__builtin_nontemporal_load(&qp[r]);
        B.Qn[i] = hsplat(-q);
      }
      B.A[i] = hsplat((half_t)1.0f);
    } else {
      B.Qn[i] = hsplat((half_t)0.0f);
      B.A[i] = hsplat((half_t)0.0f);
    }
  }
  if constexpr (use_band2<SB, MODE, M3>())
    f16_band2<SB, MODE, QK, FIX>(B, io, ring_s, ring_t, lane, best, posf, posr, Kl, &rowmin);
  else
    f16_sweep<SB, MODE, M3>(B, io, Kl, ring_s, ring_t, lane, best, posf, posr, rowmin);
}

// The renormalised form at S = 16 with the split schedule (KParams::
// renorm_split; whole reads, not Read Until): the first
// Q1 = Q - renorm_split_rows(Q) rows in front-padded bands of 64 x 8 rows
// from drift renorm_k, then bands of 64 x 16 rows at twice the adapted drift
// (both on the two-column sweep).  Each band's cells and
// renormalisation are those of a band of its width in f16_item.
template <int MODE, bool M3>
__device__ void f16_item_split(const KParams& p, const WorkItem& it, int lane, h2* ring_s, h2* ring_t,
                               h2* scratch) {
  static_assert(MODE == DTWAX_F16_RENORM, "the split schedule is the renormalised form's");
  constexpr int S1 = 8, S2 = 16;
  constexpr int BR1 = 64 * S1, BR2 = 64 * S2;
  const int32_t read = it.read;
  const int32_t Q = p.q_len[read];
  const float* __restrict__ qp = p.queries + p.q_off[read];
  const int32_t Q1 = Q - renorm_split_rows(Q);
  const int32_t nb1 = (Q1 + BR1 - 1) / BR1;
  const int32_t nb = nb1 + (Q - Q1) / BR2;
  const int32_t P = nb1 * BR1 - Q1;  // virtual rows in front of band 0

  h2 best = hsplat((half_t)__builtin_inff());
  int32_t posf = -1, posr = -1;
  float off_f = 0.0f, off_r = 0.0f;  // the subtracted amounts, added back in binary32
  h2 prev_min = hsplat((half_t)0.0f);
  h2 Kb = hsplat((half_t)p.renorm_k);  // per-lane drift (adapted band to band)

  // band b of SB rows per lane whose lane 0 starts at read row r0
  auto band = [&](auto width, int32_t b, int32_t r0) {
    constexpr int SB = decltype(width)::value;
    SweepIO io;
    io.ref = (const h2*)p.ref;
    io.prio = p.prio_by_progress != 0;
    h2* scr = scratch - it.x;  // indexed by absolute column
    io.trow = (b == 0) ? nullptr : scr;
    io.final_band = (b == nb - 1);
    io.orow = io.final_band ? nullptr : scr;
    io.x = it.x;
    io.c0 = it.c0;
    io.c1 = it.c1;
    io.top_diag = false;
    io.lrow = nullptr;
    io.lr_ntiles = io.lr_last = 0;
    io.tlow = prev_min;  // 0 for band 0 (free start)
    h2 rowmin = hsplat((half_t)__builtin_inff());
    f16_item_band<SB, MODE, M3, false, false>(qp, nullptr, r0 + lane * SB, io, Kb, ring_s, ring_t, lane, best,
                                             posf, posr, rowmin);
    // subtracted on every path to the last row: 64 lane drifts and the top row's lowering
    off_f += 64.0f * (float)Kb.x + (float)prev_min.x;
    off_r += 64.0f * (float)Kb.y + (float)prev_min.y;
    if (!io.final_band) {
      h2 m = rowmin;
#pragma unroll
      for (int sh = 32; sh >= 1; sh >>= 1) {
        h2 o = __builtin_bit_cast(h2, __shfl_xor(__builtin_bit_cast(int, m), sh, 64));
        m = hmin2(m, o);
      }
      prev_min = m;
      Kb = Kb + m * hsplat((half_t)(1.0f / 64.0f));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __threadfence_block();
  };
  for (int32_t b = 0; b < nb1; b++) band(std::integral_constant<int, S1>{}, b, b * BR1 - P);
  Kb = Kb + Kb;  // twice the rows per lane
  for (int32_t b = nb1; b < nb; b++) band(std::integral_constant<int, S2>{}, b, Q1 + (b - nb1) * BR2);

  uint64_t kf = (posf >= 0) ? make_key((float)best.x + off_f, posf) : ~0ull;
  uint64_t kr = (posr >= 0) ? make_key((float)best.y + off_r, posr) : ~0ull;
  publish_min(&p.keys[2 * (int64_t)read], wave_min_u64(kf), lane);
  publish_min(&p.keys[2 * (int64_t)read + 1], wave_min_u64(kr), lane);
}

template <int S, int MODE, bool M3, bool QK = false, bool FIX = false>
__device__ void f16_item(const KParams& p, const WorkItem& it, int lane, h2* ring_s, h2* ring_t,
                         h2* scratch) {
  const int32_t read = it.read;
  const int32_t Q = read_length<FIX>(p, read);
  const int64_t qo = read_offset<FIX>(p, read);
  const float* __restrict__ qp = p.queries + qo;
  const uint32_t* __restrict__ qw = (const uint32_t*)p.queries + qo;  // half2 query words
  constexpr int BR = 64 * S;
  const int32_t nb = (Q + BR - 1) / BR;
  const int32_t P = nb * BR - Q;  // virtual rows in front of band 0
  const h2 K = (MODE == DTWAX_F16_RENORM) ? hsplat((half_t)p.renorm_k) : hsplat((half_t)0.0f);

  const h2* trow_ext = nullptr;
  h2* orow_ext = nullptr;
  if (p.state) {
    h2* base = (h2*)p.state + (int64_t)it.slot * 2 * p.row_stride;
    if (it.flags & ITEM_RESUME) trow_ext = base + (int64_t)it.cur * p.row_stride;
    if (it.flags & ITEM_STORE) orow_ext = base + (int64_t)(1 - it.cur) * p.row_stride;
  }

  if constexpr (MODE == DTWAX_F16_RENORM && S == 16 && !QK && !FIX) {
    if (p.renorm_split && !p.state) {
      f16_item_split<MODE, M3>(p, it, lane, ring_s, ring_t, scratch);
      return;
    }
  }

  h2 best = hsplat((half_t)__builtin_inff());
  int32_t posf = -1, posr = -1;
  // F16_RENORM: every lane's first row is lowered by the lane's drift Kb
  // (folded into its cost), a per-row constant (the same for all columns),
  // so the last row ends up lowered by the sum of these on its path; that sum
  // is added back here in binary32.  A band after the first also stages its
  // top row lowered by that row's exact minimum, per strand.
  float off_f = 0.0f, off_r = 0.0f;
  h2 prev_min = hsplat((half_t)0.0f);
  h2 Kb = K;  // per-lane drift of the current band (adapted band to band)
  // Read Until: the engine's per-read drift, and the stored row's minimum,
  // by which the resumed top row is lowered; one offset for every
  // reference piece of the chunk (the pieces share the stored row), so no
  // band-to-band adaptation; the keys hold the lowered score
  const bool ru_rn = MODE == DTWAX_F16_RENORM && p.ru_renorm != nullptr;
  if (ru_rn) {
    Kb = __builtin_bit_cast(h2, p.ru_renorm[2 * (int64_t)read]);
    if (trow_ext) prev_min = __builtin_bit_cast(h2, p.ru_renorm[2 * (int64_t)read + 1]);
  }

  for (int32_t b = 0; b < nb; b++) {
    F16Band<S, MODE> B;
    const int32_t r0 = b * BR - P + lane * S;
#pragma unroll
    for (int i = 0; i < S; i++) {
      const int32_t r = r0 + i;
      if (r >= 0) {
        // queries are read once: a non-temporal load keeps them from evicting
        // the reference, which every work item re-reads, from L2
        if constexpr (FIX) {
          B.Qn[i] = -__builtin_bit_cast(h2, __builtin_nontemporal_load(&qw[r]));
        } else {
          half_t q = (half_t)__builtin_nontemporal_load(&qp[r]);
          B.Qn[i] = hsplat(-q);
        }
        B.A[i] = hsplat((half_t)1.0f);
      } else {
        B.Qn[i] = hsplat((half_t)0.0f);
        B.A[i] = hsplat((half_t)0.0f);
      }
    }
    SweepIO io;
    io.ref = (const h2*)p.ref;
    io.prio = p.prio_by_progress != 0;
    // the band scratch row is indexed by absolute column; it holds [x - 64, c1 + 384)
    h2* scr = scratch - it.x;
    io.trow = (b == 0) ? trow_ext : scr;
    io.final_band = (b == nb - 1);
    io.orow = io.final_band ? orow_ext : scr;
    io.x = it.x;
    io.c0 = it.c0;
    io.c1 = it.c1;
    io.top_diag = (b == 0) && (trow_ext != nullptr);
    if constexpr (FIX) {
      io.lrow = p.last_row ? (h2*)p.last_row + (int64_t)read * p.lr_stride : nullptr;
      io.lr_ntiles = p.lr_ntiles;
      io.lr_last = p.lr_last;
    }
    h2 rowmin = hsplat((half_t)__builtin_inff());
    // the top row's lowering: the previous band's row minimum, or (Read
    // Until, first band) the stored row's
    const h2 tlow = (MODE == DTWAX_F16_RENORM && (ru_rn ? b == 0 : b > 0)) ? prev_min : hsplat((half_t)0.0f);
    io.tlow = tlow;
    if constexpr (use_band2<S, MODE, M3>())
      f16_band2<S, MODE, QK, FIX>(B, io, ring_s, ring_t, lane, best, posf, posr, Kb, &rowmin);
    else
      f16_sweep<S, MODE, M3>(B, io, Kb, ring_s, ring_t, lane, best, posf, posr, rowmin);
    if (MODE == DTWAX_F16_RENORM) {
      // subtracted on every path to the last row: 64 lane drifts and the
      // top row's lowering
      off_f += 64.0f * (float)Kb.x + (float)tlow.x;
      off_r += 64.0f * (float)Kb.y + (float)tlow.y;
      if (!io.final_band && !ru_rn) {
        h2 m = rowmin;
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) {
          h2 o = __builtin_bit_cast(h2, __shfl_xor(__builtin_bit_cast(int, m), sh, 64));
          m = hmin2(m, o);
        }
        prev_min = m;
        // The row minimum still rose by m over this band's 64 lane hops:
        // refine the per-lane drift for the next band (per strand).
        Kb = Kb + m * hsplat((half_t)(1.0f / 64.0f));
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __threadfence_block();
  }

  // Combine: per strand min value (as float + offset), lowest column on ties.
  if (ru_rn) off_f = off_r = 0.0f;  // the engine adds the offsets back
  uint64_t kf = (posf >= 0) ? make_key((float)best.x + off_f, posf) : ~0ull;
  uint64_t kr = (posr >= 0) ? make_key((float)best.y + off_r, posr) : ~0ull;
  publish_min(&p.keys[2 * (int64_t)read], wave_min_u64(kf), lane);
  publish_min(&p.keys[2 * (int64_t)read + 1], wave_min_u64(kr), lane);
}

template <int S, int MODE, bool M3, bool QK = false, bool FIX = false>
__global__ __launch_bounds__(64) void sdtw_f16_kernel(KParams p) {
  constexpr int RA = use_band2<S, MODE, M3>() ? RING2_ALLOC : RING_ALLOC;
  __shared__ h2 ring_s[RA];
  __shared__ h2 ring_t[RA + 2 * HSTRIP];  // + the sweeps' last-row strips
  const int lane = threadIdx.x;
  h2* scratch = (h2*)p.scratch + (int64_t)blockIdx.x * p.scratch_stride + SCRATCH_PADL;
  for (;;) {
    const int idx = fetch_item(p.counter, lane);
    if (idx >= p.n_items) break;
    WorkItem it = load_item<FIX>(p, idx);
    f16_item<S, MODE, M3, QK, FIX>(p, it, lane, ring_s, ring_t, scratch);
  }
}

// A pack: the shared band (the first, partial band of each of its reads),
// the scores of the reads that end there, then each continuing read's full
// bands in order, from its scratch row, as an unpacked item's bands.
template <int S, int MODE>
__device__ void f16_pack_item(const KParams& p, const WorkItem& it, int lane, h2* ring_s, h2* ring_t,
                              h2* scratch_w) {
  constexpr int BR = 64 * S;
  const h2 INF = hsplat((half_t)__builtin_inff());
  const PackLane pl = p.pack_lanes[64 * (int64_t)it.read + lane];
  SweepIO io;
  io.ref = (const h2*)p.ref;
  io.prio = p.prio_by_progress != 0;
  io.trow = nullptr;
  io.orow = nullptr;
  io.x = it.x;
  io.c0 = it.c0;
  io.c1 = it.c1;
  io.final_band = true;
  io.top_diag = false;
  io.lrow = nullptr;  // packed plans never write the tiled last row
  io.lr_ntiles = io.lr_last = 0;
  io.tlow = hsplat((half_t)0.0f);
  h2 best = INF;
  int32_t posf = -1, posr = -1;
  {
    F16Band<S, MODE> B;
    const float* __restrict__ qp = p.queries + (pl.read >= 0 ? p.q_off[pl.read] : 0);
#pragma unroll
    for (int i = 0; i < S; i++) {
      const int32_t r = pl.r0 + i;
      if (pl.read >= 0 && r >= 0) {
        half_t q = (half_t)__builtin_nontemporal_load(&qp[r]);
        B.Qn[i] = hsplat(-q);
        B.A[i] = hsplat((half_t)1.0f);
      } else {
        B.Qn[i] = hsplat((half_t)0.0f);
        B.A[i] = hsplat((half_t)0.0f);
      }
    }
    h2* hrow = (pl.info & PL_HAND) ? scratch_w + (int64_t)pl.row * p.scratch_stride - it.x : nullptr;
    f16_band2_pack<S, MODE>(B, io, ring_s, ring_t, lane, pl.info, hrow, it.reserved, best, posf, posr);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __threadfence_block();
  {
    // every lane issues the atomics; lanes without a score pass the identity
    const bool fin = (pl.info & PL_FINAL) != 0;
    const uint64_t kf = (fin && posf >= 0) ? make_key((float)best.x, posf) : ~0ull;
    const uint64_t kr = (fin && posr >= 0) ? make_key((float)best.y, posr) : ~0ull;
    unsigned long long* kp = &p.keys[2 * (int64_t)(pl.read >= 0 ? pl.read : 0)];
    atomicMin(kp, (unsigned long long)kf);
    atomicMin(kp + 1, (unsigned long long)kr);
  }
  for (int32_t j = 0; j < it.cur; j++) {
    const PackRead pr = p.pack_reads[it.slot + j];
    const float* __restrict__ qp = p.queries + p.q_off[pr.read];
    h2* scr = scratch_w + (int64_t)pr.row * p.scratch_stride - it.x;
    // the read's own overlap: a cell that can reach a counted column
    // (r, c >= c0 - (Q - 1) + r) depends only on columns from c0 - (Q - 1) on
    io.x = max(it.x, it.c0 - (p.q_len[pr.read] - 1));
    best = INF;
    posf = posr = -1;
    for (int32_t b = 0; b < pr.nfull; b++) {
      F16Band<S, MODE> B;
      const int32_t r0 = pr.rp + b * BR + lane * S;
#pragma unroll
      for (int i = 0; i < S; i++) {
        half_t q = (half_t)__builtin_nontemporal_load(&qp[r0 + i]);
        B.Qn[i] = hsplat(-q);
        B.A[i] = hsplat((half_t)1.0f);
      }
      io.trow = scr;
      io.final_band = (b == pr.nfull - 1);
      io.orow = io.final_band ? nullptr : scr;
      f16_band2<S, MODE>(B, io, ring_s, ring_t, lane, best, posf, posr);
      __builtin_amdgcn_s_waitcnt(0);
      __threadfence_block();
    }
    const uint64_t kf = (posf >= 0) ? make_key((float)best.x, posf) : ~0ull;
    const uint64_t kr = (posr >= 0) ? make_key((float)best.y, posr) : ~0ull;
    publish_min(&p.keys[2 * (int64_t)pr.read], wave_min_u64(kf), lane);
    publish_min(&p.keys[2 * (int64_t)pr.read + 1], wave_min_u64(kr), lane);
  }
}

template <int S, int MODE>
__global__ __launch_bounds__(64) void sdtw_f16_pack_kernel(KParams p) {
  __shared__ h2 ring_s[RING2_ALLOC];
  __shared__ h2 ring_t[RING2_ALLOC + 2 * HSTRIP];  // + f16_band2's last-row strips
  const int lane = threadIdx.x;
  h2* scratch_w = p.scratch ? (h2*)p.scratch + (int64_t)blockIdx.x * p.scratch_rows * p.scratch_stride + SCRATCH_PADL
                            : nullptr;
  for (;;) {
    const int idx = fetch_item(p.counter, lane);
    if (idx >= p.n_items) break;
    WorkItem it = p.items[idx];
    f16_pack_item<S, MODE>(p, it, lane, ring_s, ring_t, scratch_w);
  }
}

// ------------------------------------------------------------------------
// binary32 form (DTWAX_F32), NSTR strands (1: single sequence, 2: fwd/rev)
// ------------------------------------------------------------------------
// One step of NSTR strands: Do = column c-1 values, Dn = column c values.
// The strands' chains (min -> add -> min ...) are independent, so they are
// interleaved cell by cell, and each strand's row costs t*t are computed two
// cells ahead of the chain; scheduling barriers keep that order (left to
// itself the scheduler emits one strand's chain as a run of dependent ops).
// Binary32 VALU ops issue at twice the rate of packed binary16 ones, so the
// dependent pairs need independent work between them to keep the issue rate.
// From S = 25 up (two waves per SIMD) the look-ahead also runs across
// steps: the last LA cells' cost slots compute the next column's (s_next)
// first LA row costs into cn, which the next step starts from, so no chain
// op runs without independent work beside it (S = 32: 13.2 -> 14.7 T
// cells/s on the 48.5 kb x 4096 x 2000 batch).  With three waves per SIMD
// (S = 16) the step's own first costs, computed at its start, hide the lane
// move better (14.7 against 14.1), so there the look-ahead stays per step.
constexpr int F32_LA = 2;  // cells of look-ahead for the row costs
template <int S>
constexpr bool f32_xstep() { return S >= 25; }
// QX: rows 2..S-1 of the lanes with xdiag set take the next lane's row i - 2
// of this step as their diagonal (the seam deviation Q2, below).
template <int S, int NSTR, bool M3, bool QX = false>
__device__ __forceinline__ void f32_cells(const float (&A)[S], const float (&Qn)[S],
                                          const float (&s)[NSTR], const float (&s_next)[NSTR],
                                          float (&cn)[NSTR][F32_LA], const float (&tin)[NSTR],
                                          const float (&dg)[NSTR], const float (&Do)[NSTR][S],
                                          float (&Dn)[NSTR][S], bool xdiag = false) {
  constexpr int LA = F32_LA;
  constexpr bool XSTEP = f32_xstep<S>();
  static_assert(S >= LA, "look-ahead within one column");
  static_assert(!QX || (NSTR == 1 && !M3), "seam deviations: one sequence, two moves");
  float tp[NSTR], dd[NSTR], c[NSTR][S];
#pragma unroll
  for (int st = 0; st < NSTR; st++) {
    tp[st] = tin[st];
    dd[st] = dg[st];
#pragma unroll
    for (int i = 0; i < LA; i++) {
      if constexpr (XSTEP) {
        c[st][i] = cn[st][i];
      } else {
        const float t = __builtin_fmaf(s[st], A[i], Qn[i]);  // == rn(s - q) for a real row
        c[st][i] = t * t;
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < S; i++) {
    float m[NSTR];
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
      if constexpr (M3) {
        m[st] = __builtin_fminf(tp[st], __builtin_fminf(dd[st], Do[st][i]));
      } else if constexpr (QX) {
        float d = dd[st];
        if (i >= 2) {
          const float nb = dpp_shl1(d, Dn[st][i - 2]);  // next lane's row i - 2, this step
          d = xdiag ? nb : d;
        }
        m[st] = __builtin_fminf(tp[st], d);
      } else {
        m[st] = __builtin_fminf(tp[st], dd[st]);
      }
    }
    float t[NSTR];
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
      // == rn(s - q) for a real row
      if (i + LA < S) t[st] = __builtin_fmaf(s[st], A[i + LA], Qn[i + LA]);
      else if (XSTEP) t[st] = __builtin_fmaf(s_next[st], A[i + LA - S], Qn[i + LA - S]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
      const float v = c[st][i] + m[st];
      Dn[st][i] = v;
      tp[st] = v;
      dd[st] = Do[st][i];
    }
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
      if (i + LA < S) c[st][i + LA] = t[st] * t[st];
      else if (XSTEP) cn[st][i + LA - S] = t[st] * t[st];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// One column step; the step's column (s_cur) and the next step's top input
// (tf_cur) were read from LDS a step earlier, and this step reads the next
// ones (slot base + 65 at most: inside the mirrored ring).
template <int S, int NSTR, bool M3>
__device__ __forceinline__ void f32_step(const float (&A)[S], const float (&Qn)[S], const float* rs,
                                         const float* rt, int u, float (&s_cur)[NSTR],
                                         float (&tf_cur)[NSTR], float (&cn)[NSTR][F32_LA],
                                         float (&tin)[NSTR], float (&dg)[NSTR], float (&hist)[NSTR],
                                         const float (&Do)[NSTR][S], float (&Dn)[NSTR][S]) {
  float s_next[NSTR], tf_next[NSTR];
#pragma unroll
  for (int st = 0; st < NSTR; st++) {
    s_next[st] = rs[(u + 1) * NSTR + st];
    tf_next[st] = rt[(u + 2) * NSTR + st];
  }
  f32_cells<S, NSTR, M3>(A, Qn, s_cur, s_next, cn, tin, dg, Do, Dn);
#pragma unroll
  for (int st = 0; st < NSTR; st++) {
    dg[st] = tin[st];
    tin[st] = dpp_shr1(tf_cur[st], Dn[st][S - 1]);
    hist[st] = dpp_shl1(Dn[st][S - 1], hist[st]);
    s_cur[st] = s_next[st];
    tf_cur[st] = tf_next[st];
  }
}

// QK: the reference kernel's seam deviations (Q1, Q2 of f16_band2) in the
// binary32 sweep of its FP32 build: 2048 rows as one band (64 lanes x 32
// rows) over the one sequence fwd | rev from column 0, one column per step,
// so lane l reaches column 832k at step 832k + l.  The deviating cells:
//   Q1  lane 32 row 0: +inf diagonal at steps 832k + 32 (k >= 1) and
//       832k + 58 (column 832k + 26, k >= 0);
//   Q2  at steps 832k + l (k >= 2): lanes 31 and 63, row 0, take their own
//       previous column's row 30; lanes 30 and 62, rows 2..31, take lane
//       31's (63's) row i - 2 of the same step (f32_cells' QX form).
// The six steps per tile (t mod 832 = 30, 31, 32, 58, 62, 63) lie in three
// 8-step groups (24, 32, 56), which run f32_step_qk; the other 101 of every
// 104 groups run the plain step.
constexpr int QK32_PERIOD = TILE_W;  // steps per tile (one column per step)

template <int S, bool QX>
__device__ __forceinline__ void f32_step_qk(const float (&A)[S], const float (&Qn)[S], const float* rs,
                                            const float* rt, int u, float (&s_cur)[1], float (&tf_cur)[1],
                                            float (&cn)[1][F32_LA], float (&tin)[1], float (&dg)[1],
                                            float (&hist)[1], const float (&Do)[1][S], float (&Dn)[1][S],
                                            int lane, int32_t t, int32_t r) {
  constexpr int NSTR = 1;
  float s_next[NSTR], tf_next[NSTR];
  s_next[0] = rs[u + 1];
  tf_next[0] = rt[u + 2];
  const bool q1 = (r == 32 && t >= QK32_PERIOD + 32) || r == 58;
  const bool q2 = (r == 30 || r == 31 || r == 62 || r == 63) && t >= 2 * QK32_PERIOD + r;
  float dq[NSTR] = {dg[0]};
  if (q1 && lane == 32) dq[0] = __builtin_inff();
  if (q2 && (r & 1) && lane == r) dq[0] = Do[0][S - 2];  // lanes 31, 63: own row 30 of column c - 1
  f32_cells<S, NSTR, false, QX>(A, Qn, s_cur, s_next, cn, tin, dq, Do, Dn, QX && q2 && lane == r);
#pragma unroll
  for (int st = 0; st < NSTR; st++) {
    dg[st] = tin[st];
    tin[st] = dpp_shr1(tf_cur[st], Dn[st][S - 1]);
    hist[st] = dpp_shl1(Dn[st][S - 1], hist[st]);
    s_cur[st] = s_next[st];
    tf_cur[st] = tf_next[st];
  }
}

template <int S, int NSTR, bool M3, bool FIX = false, bool QK = false>
__device__ void f32_item(const KParams& p, const WorkItem& it, int lane, float* ring_s,
                         float* ring_t, float* scratch) {
  static_assert(!QK || (S == 32 && NSTR == 1 && !M3 && FIX),
                "seam deviations: binary32 per op, 64 x 32 rows, one sequence from column 0");
  const int32_t read = it.read;
  const int32_t Q = read_length<FIX>(p, read);
  const float* __restrict__ qp = p.queries + read_offset<FIX>(p, read);
  constexpr int BR = 64 * S;
  const int32_t nb = (Q + BR - 1) / BR;
  const int32_t P = nb * BR - Q;
  const float INF = __builtin_inff();
  float* lrow = (FIX && p.last_row) ? (float*)p.last_row + (int64_t)read * p.lr_stride * NSTR : nullptr;
  const float* __restrict__ ref = (const float*)p.ref;  // NSTR floats per column
  const float* trow_ext = nullptr;
  float* orow_ext = nullptr;
  if (p.state) {
    float* base = (float*)p.state + (int64_t)it.slot * 2 * p.row_stride * NSTR;
    if (it.flags & ITEM_RESUME) trow_ext = base + (int64_t)it.cur * p.row_stride * NSTR;
    if (it.flags & ITEM_STORE) orow_ext = base + (int64_t)(1 - it.cur) * p.row_stride * NSTR;
  }

  float best[NSTR];
  int32_t pos[NSTR];
#pragma unroll
  for (int k = 0; k < NSTR; k++) { best[k] = INF; pos[k] = -1; }

  for (int32_t b = 0; b < nb; b++) {
    float Qn[S], A[S];
    const int32_t r0 = b * BR - P + lane * S;
#pragma unroll
    for (int i = 0; i < S; i++) {
      const int32_t r = r0 + i;
      Qn[i] = (r >= 0) ? -__builtin_nontemporal_load(&qp[r]) : 0.0f;
      A[i] = (r >= 0) ? 1.0f : 0.0f;
    }
    float* scr = scratch - (int64_t)it.x * NSTR;  // absolute-column view of the band scratch
    const float* trow = (b == 0) ? trow_ext : scr;
    const bool final_band = (b == nb - 1);
    float* orow = final_band ? orow_ext : scr;
    const int32_t x = it.x, c1 = it.c1, c0 = it.c0;
    const int32_t nsteps = ((c1 - x + 63) + 63) & ~63;

    for (int k = 0; k < 3; k++) {
      const int32_t c = x - 64 + 64 * k + lane;
#pragma unroll
      for (int st = 0; st < NSTR; st++) {
        float v = ref[(int64_t)c * NSTR + st];
        float tv = trow ? trow[(int64_t)c * NSTR + st] : 0.0f;
        ring_s[(64 * k + lane) * NSTR + st] = v;
        ring_t[(64 * k + lane) * NSTR + st] = tv;
        if (k == 0) {
          ring_s[(RING + lane) * NSTR + st] = v;
          ring_t[(RING + lane) * NSTR + st] = tv;
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();

    float D0[NSTR][S], D1[NSTR][S];
    float tin[NSTR], dg[NSTR], hist[NSTR];
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
#pragma unroll
      for (int i = 0; i < S; i++) D0[st][i] = INF;
      tin[st] = INF;
      dg[st] = INF;
      hist[st] = INF;
      if (lane == 0) tin[st] = ring_t[64 * NSTR + st];
      if (lane == 0 && b == 0 && trow_ext) dg[st] = ring_t[63 * NSTR + st];
    }

    // column of step 0 and the top input of step 1 (then read a step ahead),
    // and step 0's first row costs (the look-ahead across steps)
    float s_cur[NSTR], tf_cur[NSTR], cn[NSTR][F32_LA];
#pragma unroll
    for (int st = 0; st < NSTR; st++) {
      s_cur[st] = ring_s[((64 - lane) & (RING - 1)) * NSTR + st];
      tf_cur[st] = ring_t[(((64 - lane) & (RING - 1)) + 1) * NSTR + st];
#pragma unroll
      for (int i = 0; i < F32_LA; i++) {
        const float t = f32_xstep<S>() ? __builtin_fmaf(s_cur[st], A[i], Qn[i]) : 0.0f;
        cn[st][i] = t * t;
      }
    }
    int32_t tmod = 0;  // QK: T mod QK32_PERIOD
    for (int32_t T = 0; T < nsteps; T += 64) {
      float pf_s[NSTR], pf_t[NSTR];
#pragma unroll
      for (int st = 0; st < NSTR; st++) {
        pf_s[st] = ref[(int64_t)(x + T + 128 + lane) * NSTR + st];
        pf_t[st] = trow ? trow[(int64_t)(x + T + 128 + lane) * NSTR + st] : 0.0f;
      }
      const int base = (64 - lane + T) & (RING - 1);
#pragma unroll 1
      for (int g = 0; g < 64; g += 8) {
        const float* rs = ring_s + (base + g) * NSTR;
        const float* rt = ring_t + (base + g) * NSTR;
        if constexpr (QK) {
          // tmod is a multiple of 64 below 832 = 13 x 64, so r0 < 832
          const int32_t r0 = tmod + g;
          if (r0 == 24 || r0 == 32 || r0 == 56) {
            // a group holding a deviating step (uniform across the wave);
            // only u = 6 of groups 24 and 56 (steps 30, 62) needs the
            // cross-lane diagonal
            const int32_t t0 = T + g;
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
              if (u == 6)
                f32_step_qk<S, true>(A, Qn, rs, rt, u, s_cur, tf_cur, cn, tin, dg, hist, D0, D1, lane, t0 + u,
                                     r0 + u);
              else
                f32_step_qk<S, false>(A, Qn, rs, rt, u, s_cur, tf_cur, cn, tin, dg, hist, D0, D1, lane, t0 + u,
                                      r0 + u);
              f32_step_qk<S, false>(A, Qn, rs, rt, u + 1, s_cur, tf_cur, cn, tin, dg, hist, D1, D0, lane,
                                    t0 + u + 1, r0 + u + 1);
            }
            continue;
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          f32_step<S, NSTR, M3>(A, Qn, rs, rt, u, s_cur, tf_cur, cn, tin, dg, hist, D0, D1);
          f32_step<S, NSTR, M3>(A, Qn, rs, rt, u + 1, s_cur, tf_cur, cn, tin, dg, hist, D1, D0);
        }
      }
      const int32_t col = x + T + lane - 63;
      const bool in_item = (col >= x) && (col < c1);
#pragma unroll
      for (int st = 0; st < NSTR; st++) {
        if (final_band && col >= c0 && col < c1 && hist[st] < best[st]) {
          best[st] = hist[st];
          pos[st] = col;
        }
        if (orow && in_item && (!final_band || col >= c0)) orow[(int64_t)col * NSTR + st] = hist[st];
        if constexpr (FIX)
          if (lrow && final_band && col >= c0 && col < c1)
            lrow[tiled_index(p.lr_ntiles, p.lr_last, col) * NSTR + st] = hist[st];
      }
      __builtin_amdgcn_wave_barrier();
      const int slot = (T + 192 + lane) & (RING - 1);
#pragma unroll
      for (int st = 0; st < NSTR; st++) {
        ring_s[slot * NSTR + st] = pf_s[st];
        ring_t[slot * NSTR + st] = pf_t[st];
        if (slot < 64) {
          ring_s[(RING + slot) * NSTR + st] = pf_s[st];
          ring_t[(RING + slot) * NSTR + st] = pf_t[st];
        }
      }
      __builtin_amdgcn_wave_barrier();
      if constexpr (QK) {
        tmod += 64;
        if (tmod >= QK32_PERIOD) tmod -= QK32_PERIOD;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __threadfence_block();
  }
#pragma unroll
  for (int st = 0; st < NSTR; st++) {
    uint64_t k = (pos[st] >= 0) ? make_key(best[st], pos[st]) : ~0ull;
    publish_min(&p.keys[2 * (int64_t)read + st], wave_min_u64(k), lane);
  }
}

template <int S, int NSTR, bool M3, bool FIX = false, bool QK = false>
__global__ __launch_bounds__(64) void sdtw_f32_kernel(KParams p) {
  __shared__ float ring_s[RING_ALLOC * NSTR];
  __shared__ float ring_t[RING_ALLOC * NSTR];
  const int lane = threadIdx.x;
  float* scratch = (float*)p.scratch + ((int64_t)blockIdx.x * p.scratch_stride + SCRATCH_PADL) * NSTR;
  for (;;) {
    const int idx = fetch_item(p.counter, lane);
    if (idx >= p.n_items) break;
    WorkItem it = load_item<FIX>(p, idx);
    f32_item<S, NSTR, M3, FIX, QK>(p, it, lane, ring_s, ring_t, scratch);
  }
}

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
__global__ void reset_score_kernel(int32_t* counters, int32_t n_counters, unsigned long long* keys,
                                   int64_t n_keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_counters) counters[i] = 0;
  for (; i < n_keys; i += (int64_t)gridDim.x * blockDim.x) keys[i] = ~0ull;
}

__global__ void finalize_kernel(const unsigned long long* keys, int32_t n_reads, int32_t n_strands,
                                DeviceResult* out) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_reads) return;
  DeviceResult r;
  for (int k = 0; k < 2; k++) {
    unsigned long long key = (k < n_strands) ? keys[2 * (int64_t)i + k] : ~0ull;
    if (key == ~0ull) {
      r.score[k] = __builtin_inff();
      r.argmin[k] = -1;
    } else {
      uint32_t hi = (uint32_t)(key >> 32);
      uint32_t u = (hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi;
      r.score[k] = __builtin_bit_cast(float, u);
      r.argmin[k] = (int32_t)(key & 0xffffffffu);
    }
  }
  out[i] = r;
}

// keys -> the reference's dists layout (one half2 {fwd, rev} word or one
// float per read); a read with no column scored gets +inf.
__global__ void finalize_dists_kernel(const unsigned long long* keys, int32_t n_reads, int32_t h2out,
                                      void* dists) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_reads) return;
  float v[2];
  for (int k = 0; k < 2; k++) {
    const unsigned long long key = keys[2 * (int64_t)i + k];
    const uint32_t hi = (uint32_t)(key >> 32);
    const uint32_t u = (hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi;
    v[k] = key == ~0ull ? __builtin_inff() : __builtin_bit_cast(float, u);
  }
  if (h2out) {
    // binary16 scores are exact in float, so the conversion back is exact
    const h2 w = (h2){(half_t)v[0], (half_t)v[1]};
    ((uint32_t*)dists)[i] = __builtin_bit_cast(uint32_t, w);
  } else {
    ((float*)dists)[i] = v[0];
  }
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void selftest_lanes_kernel(int32_t* out) {
  int lane = threadIdx.x;
  out[lane] = dpp_shr1(-1, lane);
  out[64 + lane] = dpp_shl1(-1, lane);
}

// ------------------------------------------------------------------------
// launch table
// ------------------------------------------------------------------------
// m3 (three moves) exists for the exact binary16 form and binary32 only
template <int S, int MODE>
static hipError_t launch_f16(const KParams& p, int grid, bool m3, hipStream_t st) {
  if constexpr (MODE == DTWAX_F16) {
    if (m3) {
      hipLaunchKernelGGL((sdtw_f16_kernel<S, MODE, true>), dim3(grid), dim3(64), 0, st, p);
      return hipGetLastError();
    }
  } else {
    if (m3) return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL((sdtw_f16_kernel<S, MODE, false>), dim3(grid), dim3(64), 0, st, p);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_f32(const KParams& p, int grid, int nstr, bool m3, hipStream_t st) {
  if (nstr == 1) {
    if (m3) hipLaunchKernelGGL((sdtw_f32_kernel<S, 1, true>), dim3(grid), dim3(64), 0, st, p);
    else hipLaunchKernelGGL((sdtw_f32_kernel<S, 1, false>), dim3(grid), dim3(64), 0, st, p);
  } else {
    if (m3) hipLaunchKernelGGL((sdtw_f32_kernel<S, 2, true>), dim3(grid), dim3(64), 0, st, p);
    else hipLaunchKernelGGL((sdtw_f32_kernel<S, 2, false>), dim3(grid), dim3(64), 0, st, p);
  }
  return hipGetLastError();
}

template <int MODE>
static hipError_t launch_f16_s(const KParams& p, int S, int grid, bool m3, hipStream_t st) {
  switch (S) {
    case 8: return launch_f16<8, MODE>(p, grid, m3, st);
    case 12: return launch_f16<12, MODE>(p, grid, m3, st);
    case 16: return launch_f16<16, MODE>(p, grid, m3, st);
    case 20: return launch_f16<20, MODE>(p, grid, m3, st);
    case 25: return launch_f16<25, MODE>(p, grid, m3, st);
    case 28: return launch_f16<28, MODE>(p, grid, m3, st);
    case 32: return launch_f16<32, MODE>(p, grid, m3, st);
  }
  return hipErrorInvalidValue;
}

template <int MODE>
static hipError_t launch_f16_pack(const KParams& p, int S, int grid, hipStream_t st) {
  switch (S) {
    case 16: hipLaunchKernelGGL((sdtw_f16_pack_kernel<16, MODE>), dim3(grid), dim3(64), 0, st, p); break;
    case 20: hipLaunchKernelGGL((sdtw_f16_pack_kernel<20, MODE>), dim3(grid), dim3(64), 0, st, p); break;
    case 25: hipLaunchKernelGGL((sdtw_f16_pack_kernel<25, MODE>), dim3(grid), dim3(64), 0, st, p); break;
    case 28: hipLaunchKernelGGL((sdtw_f16_pack_kernel<28, MODE>), dim3(grid), dim3(64), 0, st, p); break;
    case 32: hipLaunchKernelGGL((sdtw_f16_pack_kernel<32, MODE>), dim3(grid), dim3(64), 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sdtw_fixed(const KParams& p, int precision, int grid, bool quirks, hipStream_t st) {
  if (precision == DTWAX_F16) {
    if (quirks) hipLaunchKernelGGL((sdtw_f16_kernel<32, DTWAX_F16, false, true, true>), dim3(grid), dim3(64), 0, st, p);
    else hipLaunchKernelGGL((sdtw_f16_kernel<32, DTWAX_F16, false, false, true>), dim3(grid), dim3(64), 0, st, p);
  } else if (precision == DTWAX_F32) {
    if (quirks) hipLaunchKernelGGL((sdtw_f32_kernel<32, 1, false, true, true>), dim3(grid), dim3(64), 0, st, p);
    else hipLaunchKernelGGL((sdtw_f32_kernel<32, 1, false, true>), dim3(grid), dim3(64), 0, st, p);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sdtw(const KParams& p, int precision, int S, int nstr, int grid, bool m3,
                       hipStream_t st, bool pack) {
  if (pack) {
    if (m3) return hipErrorInvalidValue;
    if (precision == DTWAX_F16) return launch_f16_pack<DTWAX_F16>(p, S, grid, st);
    if (precision == DTWAX_F16_FUSED) return launch_f16_pack<DTWAX_F16_FUSED>(p, S, grid, st);
    return hipErrorInvalidValue;
  }
  switch (precision) {
    case DTWAX_F16: return launch_f16_s<DTWAX_F16>(p, S, grid, m3, st);
    case DTWAX_F16_FUSED: return launch_f16_s<DTWAX_F16_FUSED>(p, S, grid, m3, st);
    case DTWAX_F16_RENORM: return launch_f16_s<DTWAX_F16_RENORM>(p, S, grid, m3, st);
    case DTWAX_F32:
      switch (S) {
        case 8: return launch_f32<8>(p, grid, nstr, m3, st);
        case 12: return launch_f32<12>(p, grid, nstr, m3, st);
        case 16: return launch_f32<16>(p, grid, nstr, m3, st);
        case 20: return launch_f32<20>(p, grid, nstr, m3, st);
        case 25: return launch_f32<25>(p, grid, nstr, m3, st);
        case 28: return launch_f32<28>(p, grid, nstr, m3, st);
        case 32: return launch_f32<32>(p, grid, nstr, m3, st);
      }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_reset(int32_t* counters, int32_t n_counters, unsigned long long* keys, int64_t n_keys,
                        hipStream_t st) {
  int64_t blocks = (n_keys + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(reset_score_kernel, dim3((unsigned)blocks), dim3(256), 0, st, counters, n_counters, keys,
                     n_keys);
  return hipGetLastError();
}

hipError_t launch_finalize(const unsigned long long* keys, int32_t n_reads, int32_t n_strands,
                           DeviceResult* out, hipStream_t st) {
  int blocks = (n_reads + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(256), 0, st, keys, n_reads, n_strands, out);
  return hipGetLastError();
}

hipError_t launch_finalize_dists(const unsigned long long* keys, int32_t n_reads, bool h2, void* dists,
                                 hipStream_t st) {
  if (n_reads <= 0) return hipSuccess;
  hipLaunchKernelGGL(finalize_dists_kernel, dim3((n_reads + 255) / 256), dim3(256), 0, st, keys, n_reads,
                     h2 ? 1 : 0, dists);
  return hipGetLastError();
}

hipError_t launch_selftest_lanes(int32_t* d_out, hipStream_t st) {
  hipLaunchKernelGGL(selftest_lanes_kernel, dim3(1), dim3(64), 0, st, d_out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, value);
  return hipGetLastError();
}

int occupancy_f16(int precision, int S, bool m3) {
  int n = 0;
#define OCC(MODE, SS)                                                                          \
  if (m3 && MODE == DTWAX_F16)                                                                 \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f16_kernel<SS, DTWAX_F16, true>, 64, 0); \
  else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f16_kernel<SS, MODE, false>, 64, 0);
#define OCC_S(MODE)            \
  switch (S) {                 \
    case 8: OCC(MODE, 8) break;   \
    case 12: OCC(MODE, 12) break; \
    case 16: OCC(MODE, 16) break; \
    case 20: OCC(MODE, 20) break; \
    case 25: OCC(MODE, 25) break; \
    case 28: OCC(MODE, 28) break; \
    case 32: OCC(MODE, 32) break; \
  }
  switch (precision) {
    case DTWAX_F16: OCC_S(DTWAX_F16) break;
    case DTWAX_F16_FUSED: OCC_S(DTWAX_F16_FUSED) break;
    case DTWAX_F16_RENORM: OCC_S(DTWAX_F16_RENORM) break;
  }
#undef OCC_S
#undef OCC
  return n;
}

int occupancy_fixed(int precision, bool quirks) {
  int n = 0;
  if (precision == DTWAX_F32 && quirks)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f32_kernel<32, 1, false, true, true>, 64, 0);
  else if (precision == DTWAX_F32)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f32_kernel<32, 1, false, true>, 64, 0);
  else if (quirks)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f16_kernel<32, DTWAX_F16, false, true, true>, 64, 0);
  else
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f16_kernel<32, DTWAX_F16, false, false, true>, 64, 0);
  return n;
}

int occupancy_f16_pack(int precision, int S) {
  int n = 0;
#define OCCP(MODE, SS)                                                                          \
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f16_pack_kernel<SS, MODE>, 64, 0);
#define OCCP_S(MODE)                 \
  switch (S) {                       \
    case 16: OCCP(MODE, 16) break;   \
    case 20: OCCP(MODE, 20) break;   \
    case 25: OCCP(MODE, 25) break;   \
    case 28: OCCP(MODE, 28) break;   \
    case 32: OCCP(MODE, 32) break;   \
  }
  if (precision == DTWAX_F16) OCCP_S(DTWAX_F16)
  else if (precision == DTWAX_F16_FUSED) OCCP_S(DTWAX_F16_FUSED)
#undef OCCP_S
#undef OCCP
  return n;
}

int occupancy_f32(int S, int nstr, bool m3) {
  int n = 0;
#define OCC(SS, NS)                                                                            \
  if (m3) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f32_kernel<SS, NS, true>, 64, 0); \
  else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sdtw_f32_kernel<SS, NS, false>, 64, 0);
  switch (S) {
    case 8: if (nstr == 1) { OCC(8, 1) } else { OCC(8, 2) } break;
    case 12: if (nstr == 1) { OCC(12, 1) } else { OCC(12, 2) } break;
    case 16: if (nstr == 1) { OCC(16, 1) } else { OCC(16, 2) } break;
    case 20: if (nstr == 1) { OCC(20, 1) } else { OCC(20, 2) } break;
    case 28: if (nstr == 1) { OCC(28, 1) } else { OCC(28, 2) } break;
    case 25: if (nstr == 1) { OCC(25, 1) } else { OCC(25, 2) } break;
    case 32: if (nstr == 1) { OCC(32, 1) } else { OCC(32, 2) } break;
  }
#undef OCC
  return n;
}

}  // namespace dtwax
