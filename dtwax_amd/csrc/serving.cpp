// serving.cpp -- Read Until serving engine behind the dtwax_ru_* C ABI.
//
// The reference's Read Until driver (src/ru_main.cu:172-326) scores every
// batch of 10 FIFO queries from scratch and prints the scores; it keeps no
// state and never decides.  This engine is the stateful form of the same
// computation: a read arrives as fixed-size chunks on a channel, and the
// engine keeps the read's last DP row over the whole reference (both
// strands) resident in HBM between chunks.  Chunk k then costs only its own
// rows: it resumes from the stored row, so the score after k chunks is
// bit-identical to scoring the concatenated (first-chunk-normalised) query
// from scratch with the same arithmetic.
//
// A batch = chunks of distinct channels, one persistent sDTW launch over
// (chunk x reference piece) work items.  The reference piece size is chosen
// per batch from the batch size (state rows are indexed by absolute column,
// so consecutive chunks of a read may use different pieces).
//
// Lanes: each lane is a HIP stream with its own staging buffers, work items,
// keys and results.  The serving loop launches a new batch on a free lane
// as soon as chunks arrive, while earlier batches are still running, so
// their kernels overlap on the device (a read has at most one chunk in
// flight; reads in flight hold distinct state slots).
//
// Device memory per slot: 2 rows (ping-pong: chunk k reads row cur and
// writes row 1-cur, since the pieces of one chunk run concurrently and read
// the stored row across their overlaps) x (n_cols + pads) x 4 bytes.
//
// DTWAX_F16_RENORM: every lane hop subtracts a per-row drift K (binary16,
// per strand) and lane 0 of a resumed chunk also subtracts k0, the stored
// row's exact minimum, so each chunk's values start near 0 and binary16
// keeps its resolution over any number of chunks.  A per-row constant moves
// every path's cost by the same amount, so argmins are unchanged; the
// kernel reports the lowered minimum and the engine adds the subtracted
// amounts back per slot in binary64 (64 K per band + k0 per chunk).  K
// starts at renorm_drift x segment per lane (0.01 per row by default; a
// larger start overshoots for targets, whose scores rise slowly, and their
// relative error grew) and is adapted after every chunk by the
// lowered minimum's remaining rise, K += m / 64 (the batch form's band-to-
// band rule); within a chunk it is fixed, so every reference piece of the
// chunk lowers the stored row by the same amount.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "dtwax.h"
#include "internal.h"
#include "kernels.h"

using namespace dtwax;

namespace {

using Clock = std::chrono::steady_clock;

struct ChanRec {
  int32_t number = 0;
  int32_t slot = -1;
  int32_t action = DTWAX_RU_CONTINUE;
  int32_t chunks = 0;
  float score[2] = {INFINITY, INFINITY};
  int32_t argmin[2] = {-1, -1};
};

struct Pending {
  int32_t channel;
  int32_t number;
  bool release;  // end of the channel's read (no samples)
  std::vector<float> samples;
  Clock::time_point t;
};

struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, done = nullptr;
  float* d_raw = nullptr;
  float* d_norm = nullptr;
  int32_t* d_slot = nullptr;
  int32_t* d_first = nullptr;
  WorkItem* d_items = nullptr;
  int32_t* d_counter = nullptr;
  unsigned long long* d_keys = nullptr;
  DeviceResult* d_res = nullptr;
  void* d_scratch = nullptr;
  // One pinned host block and one device block stage a batch's inputs, laid
  // out per batch (slots, first flags, renorm words, items, raw samples) so
  // that a single copy moves them; the pointers below are views into them.
  unsigned char* h_stage = nullptr;
  unsigned char* d_stage = nullptr;
  size_t stage_bytes = 0;
  uint32_t* d_rn = nullptr;  // DTWAX_F16_RENORM: {K, k0} half2 words per read (else null)
  uint32_t* h_rn = nullptr;
  float* h_raw = nullptr;
  int32_t* h_slot = nullptr;
  int32_t* h_first = nullptr;
  WorkItem* h_items = nullptr;
  DeviceResult* h_res = nullptr;
  // batch in flight
  bool busy = false;
  std::vector<int32_t> channels;
  std::vector<Clock::time_point> t_in;
};

constexpr uint32_t kHalf2Inf = 0x7c007c00u;
constexpr uint32_t kFloatInf = 0x7f800000u;

bool valid_segment(int S) {
  return S == 8 || S == 12 || S == 16 || S == 20 || S == 25 || S == 28 || S == 32;
}

}  // namespace

struct dtwax_ru {
  dtwax_ctx* ctx = nullptr;
  dtwax_ru_opts o{};
  int S = 25;
  int nstr = 2;
  int32_t slots = 0;
  int64_t n_cols = 0;      // reference columns the state rows were sized for
  int32_t n_strands = 0;
  int64_t row_stride = 0;  // columns per state row, pads included
  int64_t epc = 1;         // 4-byte elements per column
  int64_t grid_slots = 0;  // resident waves of the sDTW kernel
  int64_t n_simd = 0;      // SIMDs of the device (4 per CU)
  int64_t piece_floor = 0; // shortest piece of the one-to-three-items-per-SIMD layouts
  int64_t items_cap = 0;
  int64_t chunk_cap = 0;   // largest reference piece (columns)
  bool multiband = false;
  int64_t scratch_stride = 0;
  // shared device state
  void* d_state = nullptr;
  double* d_stats = nullptr;  // (centre, inv) per slot
  int64_t* d_qoff = nullptr;
  int32_t* d_qlen = nullptr;
  std::vector<Lane> lanes;
  // host bookkeeping (guarded by book_mu)
  std::mutex book_mu;
  std::unordered_map<int32_t, ChanRec> chans;
  std::vector<int32_t> free_slots;
  std::vector<int32_t> slot_cur;
  // DTWAX_F16_RENORM per slot: drift K and stored-row minimum k0 (binary16
  // bits per strand), and the amount subtracted so far (binary64 per strand)
  std::vector<uint16_t> rn_k, rn_k0;
  std::vector<double> rn_acc;
  int64_t batches = 0, chunks_scored = 0, cells = 0;
  double gpu_ms = 0.0;
  // serving loop
  std::thread worker;
  std::mutex mu;  // inq, outq, flags
  std::condition_variable cv_in, cv_out;
  std::deque<Pending> inq;
  std::deque<dtwax_ru_decision> outq;
  bool running = false;
  bool stop_req = false;
  int worker_rc = DTWAX_OK;
  std::string worker_err;
};

namespace {

void lane_free(Lane& l) {
  if (l.st) (void)hipStreamSynchronize(l.st);
  for (void* p : {(void*)l.d_stage, (void*)l.d_norm, (void*)l.d_counter, (void*)l.d_keys, (void*)l.d_res,
                  l.d_scratch})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)l.h_stage, (void*)l.h_res})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t e : {l.ev0, l.ev1, l.done})
    if (e) (void)hipEventDestroy(e);
  if (l.st) (void)hipStreamDestroy(l.st);
  l = Lane{};
}

constexpr size_t kStageAlign = 256;
size_t stage_part(size_t bytes) { return (bytes + kStageAlign - 1) / kStageAlign * kStageAlign; }

bool lane_alloc(dtwax_ru* r, Lane& l) {
  const size_t slots = (size_t)r->slots, L = (size_t)r->o.chunk_len;
  bool ok = hipStreamCreate(&l.st) == hipSuccess && hipEventCreate(&l.ev0) == hipSuccess &&
            hipEventCreate(&l.ev1) == hipSuccess &&
            hipEventCreateWithFlags(&l.done, hipEventDisableTiming) == hipSuccess;
  // the largest batch's stage: every part rounded up to kStageAlign
  l.stage_bytes = stage_part(slots * sizeof(int32_t)) * 2 + stage_part(slots * 2 * sizeof(uint32_t)) +
                  stage_part((size_t)r->items_cap * sizeof(WorkItem)) + stage_part(slots * L * sizeof(float));
  ok = ok && hipMalloc(&l.d_stage, l.stage_bytes) == hipSuccess &&
       hipMalloc(&l.d_norm, slots * L * sizeof(float)) == hipSuccess &&
       hipMalloc(&l.d_counter, sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&l.d_keys, slots * 2 * sizeof(unsigned long long)) == hipSuccess &&
       hipMalloc(&l.d_res, slots * sizeof(DeviceResult)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&l.h_stage, l.stage_bytes, 0) == hipSuccess &&
       hipHostMalloc((void**)&l.h_res, slots * sizeof(DeviceResult), 0) == hipSuccess;
  if (ok && r->multiband) {
    const int64_t n = r->grid_slots * r->scratch_stride * r->epc;
    ok = hipMalloc(&l.d_scratch, (size_t)n * 4) == hipSuccess &&
         launch_fill_u32((uint32_t*)l.d_scratch, n, 0u, l.st) == hipSuccess;
  }
  return ok;
}

void ru_free(dtwax_ru* r) {
  if (r->ctx) (void)hipSetDevice(r->ctx->device);
  for (Lane& l : r->lanes) lane_free(l);
  for (void* p : {r->d_state, (void*)r->d_stats, (void*)r->d_qoff, (void*)r->d_qlen})
    if (p) (void)hipFree(p);
  delete r;
}

// Reference piece for a batch of m chunks.  Every piece of a read repeats
// the Q - 1 columns before it (the overlap that makes pieces independent),
// and the batch ends with its slowest SIMD.  Candidates:
//  - about 8 work items per resident wave, no shorter than min_ref_chunk
//    (load balance for large batches);
//  - pieces that give every SIMD one, two or three items (small batches: one
//    sDTW wave alone on a SIMD runs almost as fast as three sharing it).
// The piece with the least modelled time of the slowest SIMD wins: items
// per wave x waves per SIMD x (piece + Q - 1) / throughput of that many
// waves.  Relative SIMD throughput at 1, 2, 3 waves, measured with one
// 1600-sample chunk against a 5 Mb reference: 1.0, 1.09, 1.2 (one item per
// SIMD, 4,928-column pieces: 1.38 ms; 782 items of 6,400 columns: 1.69 ms;
// 1,221 items of 4,096: 2.21 ms).  A per-SIMD layout has one item per wave
// and no slack for imbalance, so it must model 5% faster than the
// throughput layout: at 512 chunks it modelled 1.4% faster and ran 3%
// slower (454 against 441 ms); at 1..64 chunks it runs 7-22% faster.
int64_t piece_for(const dtwax_ru* r, int64_t m) {
  const int64_t N = r->n_cols;
  const int64_t L = r->o.chunk_len;
  const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(r->chunk_cap, N));
  auto fit = [&](int64_t p, int64_t floor) {
    p = std::max<int64_t>(p, floor);
    p = (p + 63) / 64 * 64;
    return std::max<int64_t>(1, std::min<int64_t>(p, cap));
  };
  const int64_t wmax = std::max<int64_t>(1, std::min<int64_t>(3, r->grid_slots / std::max<int64_t>(1, r->n_simd)));
  auto model = [&](int64_t p) {
    static const double kRate[4] = {0.0, 1.0, 1.09, 1.2};
    const int64_t items = m * ((N + p - 1) / p);
    const int64_t waves = std::min<int64_t>(items, r->grid_slots);
    const int64_t per_wave = (items + waves - 1) / waves;
    const int64_t w = std::min<int64_t>(wmax, (waves + r->n_simd - 1) / r->n_simd);
    return (double)per_wave * (double)w * (double)(std::min<int64_t>(p, N) + L - 1) / kRate[w];
  };
  int64_t best = fit((N * m + 8 * r->grid_slots - 1) / (8 * r->grid_slots), r->o.min_ref_chunk);
  double t_best = model(best);
  t_best *= 0.95;
  for (int64_t k = 1; k <= wmax; k++) {
    const int64_t p = fit((N * m + k * r->n_simd - 1) / (k * r->n_simd), r->piece_floor);
    const double t = model(p);
    if (t < t_best) {
      t_best = t;
      best = p;
    }
  }
  return best;
}

void set_decision(dtwax_ru_decision& d, int32_t channel, const ChanRec& c, int32_t action) {
  d.channel = channel;
  d.number = c.number;
  d.action = action;
  d.chunks = c.chunks;
  d.score[0] = c.score[0];
  d.score[1] = c.score[1];
  d.argmin[0] = c.argmin[0];
  d.argmin[1] = c.argmin[1];
  d.latency_us = 0.0;
}

uint16_t f16_bits(float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
float f16_value(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

void release_slot(dtwax_ru* r, ChanRec& c) {
  if (c.slot >= 0) r->free_slots.push_back(c.slot);
  c.slot = -1;
}

void release_channel(dtwax_ru* r, int32_t channel) {
  auto it = r->chans.find(channel);
  if (it != r->chans.end()) {
    release_slot(r, it->second);
    r->chans.erase(it);
  }
}

// Admission of one chunk (book_mu held): true if it is to be scored (the
// read holds a slot), false if it is answered at once (NO_SLOT or STALE).
bool admit(dtwax_ru* r, int32_t channel, int32_t number, dtwax_ru_decision& d) {
  auto it = r->chans.find(channel);
  if (it == r->chans.end() || it->second.number != number) {
    if (it != r->chans.end()) release_slot(r, it->second);
    ChanRec fresh;
    fresh.number = number;
    if (r->free_slots.empty()) {
      fresh.action = DTWAX_RU_NO_SLOT;
      r->chans[channel] = fresh;
      set_decision(d, channel, fresh, DTWAX_RU_NO_SLOT);
      return false;
    }
    fresh.slot = r->free_slots.back();
    r->free_slots.pop_back();
    r->chans[channel] = fresh;
    return true;
  }
  if (it->second.action != DTWAX_RU_CONTINUE) {
    set_decision(d, channel, it->second, DTWAX_RU_STALE);
    return false;
  }
  return true;
}

// Stages admitted chunks (channels[j], samples src[j]) and launches the
// batch asynchronously on lane l (book_mu held).
int launch(dtwax_ru* r, Lane& l, const std::vector<int32_t>& channels, const std::vector<const float*>& src) {
  dtwax_ctx* c = r->ctx;
  if (c->n_cols != r->n_cols || c->n_strands != r->n_strands)
    return set_error(DTWAX_ENOTREADY, "the context's reference changed after dtwax_ru_create");
  const int32_t L = r->o.chunk_len;
  const int64_t m = (int64_t)channels.size();
  const int64_t N = r->n_cols;
  const int64_t piece = piece_for(r, m);
  const int64_t per_read = (N + piece - 1) / piece;
  if (m > r->slots || m * per_read > r->items_cap) return set_error(DTWAX_EINVAL, "batch capacity exceeded");
  // this batch's stage layout; the views point into the lane's two blocks
  const bool rn = r->o.precision == DTWAX_F16_RENORM;
  size_t off = 0;
  auto part = [&](size_t bytes) {
    const size_t o = off;
    off += stage_part(bytes);
    return o;
  };
  const size_t o_slot = part((size_t)m * sizeof(int32_t));
  const size_t o_first = part((size_t)m * sizeof(int32_t));
  const size_t o_rn = part(rn ? (size_t)m * 2 * sizeof(uint32_t) : 0);
  const size_t o_items = part((size_t)(m * per_read) * sizeof(WorkItem));
  const size_t o_raw = part((size_t)(m * L) * sizeof(float));
  if (off > l.stage_bytes) return set_error(DTWAX_EINVAL, "batch stage exceeds its buffer");
  l.h_slot = (int32_t*)(l.h_stage + o_slot);
  l.h_first = (int32_t*)(l.h_stage + o_first);
  l.h_rn = rn ? (uint32_t*)(l.h_stage + o_rn) : nullptr;
  l.h_items = (WorkItem*)(l.h_stage + o_items);
  l.h_raw = (float*)(l.h_stage + o_raw);
  l.d_slot = (int32_t*)(l.d_stage + o_slot);
  l.d_first = (int32_t*)(l.d_stage + o_first);
  l.d_rn = rn ? (uint32_t*)(l.d_stage + o_rn) : nullptr;
  l.d_items = (WorkItem*)(l.d_stage + o_items);
  l.d_raw = (float*)(l.d_stage + o_raw);
  int64_t ni = 0;
  for (int64_t j = 0; j < m; j++) {
    const ChanRec& rec = r->chans[channels[(size_t)j]];
    std::memcpy(l.h_raw + j * L, src[(size_t)j], (size_t)L * sizeof(float));
    l.h_slot[j] = rec.slot;
    const bool first = rec.chunks == 0;
    l.h_first[j] = first ? 1 : 0;
    if (r->o.precision == DTWAX_F16_RENORM) {
      const size_t s2 = 2 * (size_t)rec.slot;
      if (first) {  // a new read in this slot
        const uint16_t k = f16_bits(r->o.renorm_drift * (float)r->S);
        r->rn_k[s2] = r->rn_k[s2 + 1] = k;
        r->rn_k0[s2] = r->rn_k0[s2 + 1] = 0;
        r->rn_acc[s2] = r->rn_acc[s2 + 1] = 0.0;
      }
      l.h_rn[2 * j] = (uint32_t)r->rn_k[s2] | ((uint32_t)r->rn_k[s2 + 1] << 16);
      l.h_rn[2 * j + 1] = (uint32_t)r->rn_k0[s2] | ((uint32_t)r->rn_k0[s2 + 1] << 16);
    }
    const bool store = rec.chunks + 1 < r->o.max_chunks;
    for (int64_t c0 = 0; c0 < N; c0 += piece) {
      WorkItem w{};
      w.read = (int32_t)j;
      w.slot = rec.slot;
      w.c0 = (int32_t)c0;
      w.c1 = (int32_t)std::min<int64_t>(N, c0 + piece);
      w.x = (int32_t)std::max<int64_t>(0, c0 - (L - 1));
      w.flags = (first ? 0 : ITEM_RESUME) | (store ? ITEM_STORE : 0);
      w.cur = r->slot_cur[(size_t)rec.slot];
      l.h_items[ni++] = w;
    }
  }
  DTWAX_HIP_TRY(hipSetDevice(c->device));
  // one copy for the whole stage (the raw samples last: it ends with them)
  DTWAX_HIP_TRY(hipMemcpyAsync(l.d_stage, l.h_stage, off, hipMemcpyHostToDevice, l.st));
  DTWAX_HIP_TRY(launch_ru_normalize(l.d_raw, (int32_t)m, L, l.d_slot, l.d_first, r->o.norm, r->o.eps,
                                    r->d_stats, l.d_norm, l.st));
  ScoreLaunch sl;
  sl.precision = r->o.precision;
  sl.S = r->S;
  sl.nstr = r->nstr;
  sl.grid = (int)std::min<int64_t>(r->grid_slots, ni);
  sl.queries = l.d_norm;
  sl.q_off = r->d_qoff;
  sl.q_len = r->d_qlen;
  sl.items = l.d_items;
  sl.n_items = (int32_t)ni;
  sl.n_reads = (int32_t)m;
  sl.counter = l.d_counter;
  sl.keys = l.d_keys;
  sl.scratch = l.d_scratch;
  sl.scratch_stride = r->scratch_stride;
  sl.state = (char*)r->d_state + (int64_t)COL_PADL * r->epc * 4;  // column 0 of slot 0, row 0
  sl.row_stride = r->row_stride;
  sl.ru_renorm = l.d_rn;
  sl.renorm_k = r->o.renorm_drift * (float)r->S;
  sl.out = l.d_res;
  sl.ev0 = l.ev0;
  sl.ev1 = l.ev1;
  sl.prio_ok = false;  // batches run side by side: default wave order
  int rc = run_score_launch(c, sl, l.st);
  if (rc) return rc;
  DTWAX_HIP_TRY(hipMemcpyAsync(l.h_res, l.d_res, (size_t)m * sizeof(DeviceResult), hipMemcpyDeviceToHost, l.st));
  DTWAX_HIP_TRY(hipEventRecord(l.done, l.st));
  l.channels = channels;
  l.busy = true;
  return DTWAX_OK;
}

// Waits for lane l's batch and makes the decisions, out[j] for channels[j]
// (book_mu held).
int retire(dtwax_ru* r, Lane& l, dtwax_ru_decision* out) {
  const int32_t L = r->o.chunk_len;
  DTWAX_HIP_TRY(hipEventSynchronize(l.done));
  l.busy = false;
  const int64_t m = (int64_t)l.channels.size();
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, l.ev0, l.ev1) == hipSuccess) r->gpu_ms += ms;
  r->batches++;
  r->chunks_scored += m;
  r->cells += m * (int64_t)L * r->n_cols * r->nstr;
  for (int64_t j = 0; j < m; j++) {
    const int32_t ch = l.channels[(size_t)j];
    ChanRec& rec = r->chans[ch];
    const DeviceResult& res = l.h_res[j];
    rec.chunks++;
    for (int k = 0; k < 2; k++) {
      rec.score[k] = (k < r->nstr) ? res.score[k] : INFINITY;
      rec.argmin[k] = (k < r->nstr) ? res.argmin[k] : -1;
    }
    if (r->o.precision == DTWAX_F16_RENORM) {
      // add back what this chunk subtracted (64 K per band, k0 at lane 0 of
      // a resumed chunk) and the earlier chunks' amounts; then the next
      // chunk's k0 (this row's lowered minimum) and adapted drift
      const size_t s2 = 2 * (size_t)rec.slot;
      const int64_t bands = (int64_t)L / (64 * r->S);
      for (int k = 0; k < r->nstr; k++) {
        const double K = f16_value(r->rn_k[s2 + k]);
        r->rn_acc[s2 + k] += (double)bands * 64.0 * K + (double)f16_value(r->rn_k0[s2 + k]);
        const float lowered = res.score[k];
        rec.score[k] = std::isfinite(lowered) ? (float)((double)lowered + r->rn_acc[s2 + k]) : lowered;
        if (std::isfinite(lowered)) {
          r->rn_k0[s2 + k] = f16_bits(lowered);  // a binary16 value already
          r->rn_k[s2 + k] = f16_bits((float)K + lowered / 64.0f);
        } else {
          r->rn_k0[s2 + k] = 0;
        }
      }
    }
    const float best = std::min(rec.score[0], rec.score[1]);
    const float limit = r->o.threshold_per_row > 0.0f
                            ? r->o.threshold_per_row * (float)((int64_t)rec.chunks * L)
                            : r->o.threshold;
    int32_t action = DTWAX_RU_CONTINUE;
    if (best >= limit) action = DTWAX_RU_NONTARGET;
    else if (rec.chunks >= r->o.max_chunks) action = DTWAX_RU_TARGET;
    if (action == DTWAX_RU_CONTINUE) {
      r->slot_cur[(size_t)rec.slot] ^= 1;  // the new row is in the other half
    } else {
      rec.action = action;
      release_slot(r, rec);
    }
    set_decision(out[j], ch, rec, action);
  }
  return DTWAX_OK;
}

void worker_fail(dtwax_ru* r, int rc) {
  std::lock_guard<std::mutex> lk(r->mu);
  r->worker_rc = rc;
  r->worker_err = dtwax_last_error();
  r->running = false;
  r->cv_out.notify_all();
}

void worker_main(dtwax_ru* r) {
  std::unordered_set<int32_t> inflight;  // channels with a chunk on the GPU
  std::vector<dtwax_ru_decision> dec;
  for (;;) {
    bool progressed = false;
    // 1. retire finished batches
    for (Lane& l : r->lanes) {
      if (!l.busy) continue;
      const hipError_t q = hipEventQuery(l.done);
      if (q == hipErrorNotReady) continue;
      dec.assign(l.channels.size(), dtwax_ru_decision{});
      int rc;
      {
        std::lock_guard<std::mutex> bk(r->book_mu);
        rc = (q == hipSuccess) ? retire(r, l, dec.data())
                               : set_error(DTWAX_EDEVICE, std::string("batch failed: ") + hipGetErrorString(q));
      }
      if (rc) return worker_fail(r, rc);
      const Clock::time_point now = Clock::now();
      std::lock_guard<std::mutex> lk(r->mu);
      for (size_t j = 0; j < dec.size(); j++) {
        dec[j].latency_us = std::chrono::duration<double, std::micro>(now - l.t_in[j]).count();
        inflight.erase(l.channels[j]);
        r->outq.push_back(dec[j]);
      }
      r->cv_out.notify_all();
      progressed = true;
    }
    // 2. launch what has arrived on a free lane: the earliest chunk of each
    //    channel without a chunk in flight; later ones keep their order
    Lane* free_lane = nullptr;
    for (Lane& l : r->lanes)
      if (!l.busy) { free_lane = &l; break; }
    std::vector<Pending> take;
    bool idle_exit = false;
    {
      std::unique_lock<std::mutex> lk(r->mu);
      bool any_busy = false;
      for (Lane& l : r->lanes) any_busy |= l.busy;
      if (free_lane && !r->inq.empty()) {
        std::unordered_set<int32_t> seen;
        std::deque<Pending> keep;
        while (!r->inq.empty()) {
          Pending p = std::move(r->inq.front());
          r->inq.pop_front();
          if (inflight.count(p.channel) || !seen.insert(p.channel).second) keep.push_back(std::move(p));
          else take.push_back(std::move(p));
        }
        r->inq.swap(keep);
      }
      if (take.empty() && !progressed) {
        if (r->stop_req && r->inq.empty() && !any_busy) idle_exit = true;
        else if (!any_busy && r->inq.empty()) r->cv_in.wait(lk, [&] { return r->stop_req || !r->inq.empty(); });
        // a batch is on the GPU: look again soon (a 10 us poll instead of
        // 50 us took ~25 us off the decision p50 on a 512-channel flow cell)
        else r->cv_in.wait_for(lk, std::chrono::microseconds(10));
      }
    }
    if (idle_exit) return;
    if (take.empty()) continue;
    // admission; releases and answered chunks go straight out
    std::vector<int32_t> channels;
    std::vector<const float*> src;
    std::vector<Clock::time_point> t_in;
    std::vector<dtwax_ru_decision> answered;
    int rc = DTWAX_OK;
    {
      std::lock_guard<std::mutex> bk(r->book_mu);
      for (Pending& p : take) {
        if (p.release) {
          release_channel(r, p.channel);
          continue;
        }
        dtwax_ru_decision d{};
        if (admit(r, p.channel, p.number, d)) {
          channels.push_back(p.channel);
          src.push_back(p.samples.data());
          t_in.push_back(p.t);
        } else {
          d.latency_us = std::chrono::duration<double, std::micro>(Clock::now() - p.t).count();
          answered.push_back(d);
        }
      }
      if (!channels.empty()) rc = launch(r, *free_lane, channels, src);
    }
    if (rc) return worker_fail(r, rc);
    std::lock_guard<std::mutex> lk(r->mu);
    if (!channels.empty()) {
      free_lane->t_in = std::move(t_in);
      for (int32_t ch : channels) inflight.insert(ch);
    }
    for (auto& d : answered) r->outq.push_back(d);
    if (!answered.empty()) r->cv_out.notify_all();
  }
}

bool loop_running(dtwax_ru* r) {
  std::lock_guard<std::mutex> lk(r->mu);
  return r->running;
}

}  // namespace

extern "C" {

int dtwax_ru_create(dtwax_ctx* c, const dtwax_ru_opts* opts, dtwax_ru** out) {
  if (!c || !out) return set_error(DTWAX_EINVAL, "NULL argument");
  if (!c->n_cols) return set_error(DTWAX_ENOTREADY, "no reference uploaded");
  dtwax_ru_opts o{};
  if (opts) o = *opts;
  if (o.max_reads == 0) o.max_reads = 512;
  if (o.chunk_len == 0) o.chunk_len = 1600;
  if (o.max_chunks == 0) o.max_chunks = 4;
  if (o.n_streams == 0) o.n_streams = 2;
  if (o.max_reads < 1 || o.chunk_len < 64 || o.max_chunks < 1 || o.n_streams < 1 || o.n_streams > 8)
    return set_error(DTWAX_EINVAL, "max_reads, chunk_len (>= 64), max_chunks and n_streams (1..8) "
                                   "must be positive");
  if (o.precision != DTWAX_F16 && o.precision != DTWAX_F32 && o.precision != DTWAX_F16_RENORM &&
      o.precision != DTWAX_F16_FUSED)
    return set_error(DTWAX_EINVAL, "Read Until serving supports DTWAX_F16, DTWAX_F16_FUSED, DTWAX_F16_RENORM "
                                   "and DTWAX_F32");
  if (o.norm < DTWAX_NORM_NONE || o.norm > DTWAX_NORM_MEDMAD) return set_error(DTWAX_EINVAL, "bad norm");
  if (o.chunk_len % 64) return set_error(DTWAX_EINVAL, "chunk_len must be a multiple of 64");
  const int rows_per_lane = o.chunk_len / 64;
  if (o.segment == 0) {
    for (int S : {32, 28, 25, 20, 16, 12, 8})
      if (rows_per_lane % S == 0) { o.segment = S; break; }
    if (o.segment == 0)
      return set_error(DTWAX_EINVAL, "chunk_len / 64 must be a multiple of 8, 12, 16, 20, 25, 28 or 32");
  }
  if (!valid_segment(o.segment) || rows_per_lane % o.segment)
    return set_error(DTWAX_EINVAL, "segment must be 8, 12, 16, 20, 25, 28 or 32 and divide chunk_len / 64");
  if (o.min_ref_chunk < 0 || o.hbm_budget < 0) return set_error(DTWAX_EINVAL, "negative size");
  // a given min_ref_chunk bounds every layout; by default the per-SIMD
  // layouts of small batches may go down to 2 x chunk_len
  const int64_t piece_floor = o.min_ref_chunk ? o.min_ref_chunk : 2 * (int64_t)o.chunk_len;
  if (o.min_ref_chunk == 0) o.min_ref_chunk = 4 * o.chunk_len;
  if (!(o.threshold > 0.0f)) o.threshold = INFINITY;  // no early non-target decision
  if (!(o.renorm_drift > 0.0f)) o.renorm_drift = 0.01f;

  auto* r = new dtwax_ru;
  r->ctx = c;
  r->o = o;
  r->S = o.segment;
  r->nstr = c->n_strands;
  r->epc = (o.precision == DTWAX_F32) ? c->n_strands : 1;
  r->n_cols = c->n_cols;
  r->n_strands = c->n_strands;
  r->row_stride = COL_PADL + c->n_cols + COL_PADR;
  const int64_t slot_bytes = 2 * r->row_stride * r->epc * 4;
  int64_t slots = o.max_reads;
  if (o.hbm_budget) slots = std::min<int64_t>(slots, o.hbm_budget / slot_bytes);
  if (slots < 1) {
    delete r;
    return set_error(DTWAX_ENOMEM, "hbm_budget holds no read state");
  }
  r->slots = (int32_t)slots;
  r->multiband = o.chunk_len > 64 * o.segment;
  auto cleanup = [&](int code, const std::string& m) {
    ru_free(r);
    return set_error(code, m);
  };
  if (hipSetDevice(c->device) != hipSuccess) return cleanup(DTWAX_EDEVICE, "hipSetDevice");
  r->grid_slots = (int64_t)std::max(1, occupancy(o.precision, r->S, r->nstr, false)) * c->n_cu;
  r->n_simd = 4 * (int64_t)c->n_cu;
  r->piece_floor = piece_floor;
  const int64_t N = c->n_cols;
  const int64_t L = o.chunk_len;
  // band scratch only for chunks taller than one band: cap the piece there
  r->chunk_cap = r->multiband ? std::min<int64_t>(N, std::max<int64_t>(o.min_ref_chunk, 65536)) : N;
  const int64_t min_piece = std::min<int64_t>(N, std::max<int64_t>(64, (int64_t)o.min_ref_chunk / 64 * 64));
  // the throughput layout's items, or at most wmax items per SIMD (+ one
  // rounding item per read) for the per-SIMD layouts
  r->items_cap = slots * ((N + min_piece - 1) / min_piece) + 3 * r->n_simd + slots;
  if (r->multiband) r->scratch_stride = COL_PADL + r->chunk_cap + L + COL_PADR;
  bool ok = hipMalloc(&r->d_state, (size_t)(slots * slot_bytes)) == hipSuccess &&
            hipMalloc(&r->d_stats, (size_t)slots * 2 * sizeof(double)) == hipSuccess &&
            hipMalloc(&r->d_qoff, (size_t)slots * sizeof(int64_t)) == hipSuccess &&
            hipMalloc(&r->d_qlen, (size_t)slots * sizeof(int32_t)) == hipSuccess;
  if (!ok) return cleanup(DTWAX_ENOMEM, "device allocation failed (state " + std::to_string(slots * slot_bytes) + " bytes)");
  r->lanes.resize((size_t)o.n_streams);
  for (Lane& l : r->lanes)
    if (!lane_alloc(r, l)) return cleanup(DTWAX_ENOMEM, "lane allocation failed");
  // state rows start at +inf everywhere: the left pad is column -1 of every
  // stored row (+inf by the recurrence), the rest is overwritten before use
  if (launch_fill_u32((uint32_t*)r->d_state, slots * slot_bytes / 4,
                      o.precision == DTWAX_F32 ? kFloatInf : kHalf2Inf, r->lanes[0].st) != hipSuccess)
    return cleanup(DTWAX_EDEVICE, "state fill failed");
  {
    std::vector<int64_t> off((size_t)slots);
    std::vector<int32_t> len((size_t)slots, (int32_t)L);
    for (int64_t j = 0; j < slots; j++) off[(size_t)j] = j * L;
    if (hipMemcpy(r->d_qoff, off.data(), off.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_qlen, len.data(), len.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
      return cleanup(DTWAX_EDEVICE, "initial upload failed");
  }
  r->free_slots.reserve((size_t)slots);
  for (int64_t s = slots - 1; s >= 0; s--) r->free_slots.push_back((int32_t)s);
  r->slot_cur.assign((size_t)slots, 0);
  if (o.precision == DTWAX_F16_RENORM) {
    r->rn_k.assign(2 * (size_t)slots, 0);
    r->rn_k0.assign(2 * (size_t)slots, 0);
    r->rn_acc.assign(2 * (size_t)slots, 0.0);
  }
  *out = r;
  return DTWAX_OK;
}

int dtwax_ru_destroy(dtwax_ru* r) {
  if (!r) return DTWAX_OK;
  (void)dtwax_ru_stop(r);
  ru_free(r);
  return DTWAX_OK;
}

int dtwax_ru_info(dtwax_ru* r, int32_t* slots, int32_t* active, int32_t* segment, int64_t* state_bytes) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> bk(r->book_mu);
  if (slots) *slots = r->slots;
  if (active) *active = r->slots - (int32_t)r->free_slots.size();
  if (segment) *segment = r->S;
  if (state_bytes) *state_bytes = (int64_t)r->slots * 2 * r->row_stride * r->epc * 4;
  return DTWAX_OK;
}

int dtwax_ru_stats(dtwax_ru* r, int64_t* batches, int64_t* chunks, int64_t* cells, double* gpu_ms) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> bk(r->book_mu);
  if (batches) *batches = r->batches;
  if (chunks) *chunks = r->chunks_scored;
  if (cells) *cells = r->cells;
  if (gpu_ms) *gpu_ms = r->gpu_ms;
  return DTWAX_OK;
}

int dtwax_ru_process(dtwax_ru* r, const dtwax_ru_chunk* chunks, int32_t n, const float* samples,
                     dtwax_ru_decision* out) {
  if (!r || n < 0 || (n && (!chunks || !samples || !out))) return set_error(DTWAX_EINVAL, "NULL argument");
  if (loop_running(r)) return set_error(DTWAX_EINVAL, "the serving loop is running: use dtwax_ru_submit");
  {
    std::unordered_set<int32_t> seen;
    for (int32_t i = 0; i < n; i++) {
      if (chunks[i].offset < 0 || chunks[i].channel < 0)
        return set_error(DTWAX_EINVAL, "negative channel or offset");
      if (!seen.insert(chunks[i].channel).second)
        return set_error(DTWAX_EINVAL, "two chunks of one channel in one batch");
    }
  }
  if (r->ctx->n_cols != r->n_cols || r->ctx->n_strands != r->n_strands)
    return set_error(DTWAX_ENOTREADY, "the context's reference changed after dtwax_ru_create");
  std::lock_guard<std::mutex> bk(r->book_mu);
  std::vector<int32_t> idx, channels;
  std::vector<const float*> src;
  for (int32_t i = 0; i < n; i++) {
    if (admit(r, chunks[i].channel, chunks[i].number, out[i])) {
      idx.push_back(i);
      channels.push_back(chunks[i].channel);
      src.push_back(samples + chunks[i].offset);
    }
  }
  if (channels.empty()) return DTWAX_OK;
  Lane& l = r->lanes[0];
  int rc = launch(r, l, channels, src);
  if (rc) return rc;
  std::vector<dtwax_ru_decision> dec(channels.size());
  rc = retire(r, l, dec.data());
  if (rc) return rc;
  for (size_t j = 0; j < idx.size(); j++) out[idx[j]] = dec[j];
  return DTWAX_OK;
}

int dtwax_ru_release(dtwax_ru* r, int32_t channel) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  {
    std::lock_guard<std::mutex> lk(r->mu);
    if (r->running) {  // in order with the channel's queued chunks
      Pending p;
      p.channel = channel;
      p.number = 0;
      p.release = true;
      p.t = Clock::now();
      r->inq.push_back(std::move(p));
      r->cv_in.notify_one();
      return DTWAX_OK;
    }
  }
  std::lock_guard<std::mutex> bk(r->book_mu);
  release_channel(r, channel);
  return DTWAX_OK;
}

int dtwax_ru_start(dtwax_ru* r) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(r->mu);
  if (r->running) return set_error(DTWAX_EINVAL, "serving loop already running");
  if (r->worker.joinable()) r->worker.join();
  r->stop_req = false;
  r->worker_rc = DTWAX_OK;
  r->running = true;
  r->worker = std::thread(worker_main, r);
  return DTWAX_OK;
}

int dtwax_ru_submit(dtwax_ru* r, int32_t channel, int32_t number, const float* samples) {
  if (!r || !samples || channel < 0) return set_error(DTWAX_EINVAL, "bad argument");
  Pending p;
  p.channel = channel;
  p.number = number;
  p.release = false;
  p.samples.assign(samples, samples + r->o.chunk_len);
  p.t = Clock::now();
  std::lock_guard<std::mutex> lk(r->mu);
  if (!r->running)
    return set_error(r->worker_rc ? r->worker_rc : DTWAX_ENOTREADY,
                     r->worker_rc ? "serving loop failed: " + r->worker_err : "serving loop not started");
  r->inq.push_back(std::move(p));
  r->cv_in.notify_one();
  return DTWAX_OK;
}

int dtwax_ru_poll(dtwax_ru* r, dtwax_ru_decision* out, int32_t cap, int32_t* n_out, int32_t timeout_ms) {
  if (!r || !n_out || cap < 0 || (cap && !out)) return set_error(DTWAX_EINVAL, "bad argument");
  std::unique_lock<std::mutex> lk(r->mu);
  if (r->outq.empty() && timeout_ms != 0 && r->running) {
    auto pred = [&] { return !r->outq.empty() || !r->running; };
    if (timeout_ms < 0) r->cv_out.wait(lk, pred);
    else r->cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
  }
  int32_t k = 0;
  while (k < cap && !r->outq.empty()) {
    out[k++] = r->outq.front();
    r->outq.pop_front();
  }
  *n_out = k;
  if (k == 0 && r->worker_rc != DTWAX_OK) return set_error(r->worker_rc, "serving loop failed: " + r->worker_err);
  return DTWAX_OK;
}

int dtwax_ru_stop(dtwax_ru* r) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  {
    std::lock_guard<std::mutex> lk(r->mu);
    r->stop_req = true;
    r->cv_in.notify_all();
  }
  if (r->worker.joinable()) r->worker.join();
  std::lock_guard<std::mutex> lk(r->mu);
  r->running = false;
  if (r->worker_rc != DTWAX_OK) return set_error(r->worker_rc, "serving loop failed: " + r->worker_err);
  return DTWAX_OK;
}

}  // extern "C"
