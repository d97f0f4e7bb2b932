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
// One batch = every chunk that is ready, one persistent sDTW launch over
// (chunk x reference piece) work items.  The reference piece size is chosen
// per batch from the batch size (state rows are indexed by absolute column,
// so consecutive chunks of a read may use different pieces).
//
// Device memory per slot: 2 rows (ping-pong: chunk k reads row cur and
// writes row 1-cur, since the pieces of one chunk run concurrently and read
// the stored row across their overlaps) x (n_cols + pads) x 4 bytes.

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
  std::vector<float> samples;
  Clock::time_point t;
};

constexpr uint32_t kHalf2Inf = 0x7c007c00u;
constexpr uint32_t kFloatInf = 0x7f800000u;

bool valid_segment(int S) { return S == 8 || S == 16 || S == 25 || S == 32; }

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
  int64_t items_cap = 0;
  int64_t chunk_cap = 0;   // largest reference piece (columns)
  bool multiband = false;
  // device
  void* d_state = nullptr;
  double* d_stats = nullptr;
  float* d_raw = nullptr;
  float* d_norm = nullptr;
  int64_t* d_qoff = nullptr;
  int32_t* d_qlen = nullptr;
  int32_t* d_slot = nullptr;
  int32_t* d_first = nullptr;
  WorkItem* d_items = nullptr;
  int32_t* d_counter = nullptr;
  unsigned long long* d_keys = nullptr;
  DeviceResult* d_res = nullptr;
  void* d_scratch = nullptr;
  int64_t scratch_stride = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // pinned host staging
  float* h_raw = nullptr;
  int32_t* h_slot = nullptr;
  int32_t* h_first = nullptr;
  WorkItem* h_items = nullptr;
  DeviceResult* h_res = nullptr;
  // host state
  std::unordered_map<int32_t, ChanRec> chans;
  std::vector<int32_t> free_slots;
  std::vector<int32_t> slot_cur;
  int64_t batches = 0, chunks_scored = 0, cells = 0;
  double gpu_ms = 0.0;
  std::mutex process_mu;
  // serving loop
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<Pending> inq;
  std::deque<dtwax_ru_decision> outq;
  bool running = false;
  bool stop_req = false;
  int worker_rc = DTWAX_OK;
  std::string worker_err;
};

namespace {

void ru_free(dtwax_ru* r) {
  if (r->ctx) (void)hipSetDevice(r->ctx->device);
  if (r->st) (void)hipStreamSynchronize(r->st);
  for (void* p : {r->d_state, (void*)r->d_stats, (void*)r->d_raw, (void*)r->d_norm, (void*)r->d_qoff,
                  (void*)r->d_qlen, (void*)r->d_slot, (void*)r->d_first, (void*)r->d_items,
                  (void*)r->d_counter, (void*)r->d_keys, (void*)r->d_res, r->d_scratch})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)r->h_raw, (void*)r->h_slot, (void*)r->h_first, (void*)r->h_items, (void*)r->h_res})
    if (p) (void)hipHostFree(p);
  if (r->ev0) (void)hipEventDestroy(r->ev0);
  if (r->ev1) (void)hipEventDestroy(r->ev1);
  if (r->st) (void)hipStreamDestroy(r->st);
  delete r;
}

// Reference piece for a batch of m chunks: about 8 work items per resident
// wave, never shorter than min_ref_chunk (the Q-1 overlap costs Q/piece).
int64_t piece_for(const dtwax_ru* r, int64_t m) {
  const int64_t N = r->ctx->n_cols;
  const int64_t want_items = 8 * r->grid_slots;
  int64_t piece = (N * m + want_items - 1) / want_items;
  piece = std::max<int64_t>(piece, r->o.min_ref_chunk);
  piece = std::min<int64_t>(piece, r->chunk_cap);
  piece = (piece + 63) / 64 * 64;
  return std::max<int64_t>(1, std::min<int64_t>(piece, N));
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

void release_slot(dtwax_ru* r, ChanRec& c) {
  if (c.slot >= 0) r->free_slots.push_back(c.slot);
  c.slot = -1;
}

// One batch; the caller holds process_mu.
int process_locked(dtwax_ru* r, const dtwax_ru_chunk* chunks, int32_t n, const float* samples,
                   dtwax_ru_decision* out) {
  dtwax_ctx* c = r->ctx;
  const int32_t L = r->o.chunk_len;
  if (c->n_cols != r->n_cols || c->n_strands != r->n_strands)
    return set_error(DTWAX_ENOTREADY, "the context's reference changed after dtwax_ru_create");
  // pass 1: admission and channel bookkeeping
  std::vector<int32_t> scored;  // indices into chunks
  scored.reserve((size_t)n);
  {
    std::unordered_map<int32_t, int32_t> seen;
    for (int32_t i = 0; i < n; i++)
      if (!seen.emplace(chunks[i].channel, i).second)
        return set_error(DTWAX_EINVAL, "two chunks of one channel in one batch");
  }
  for (int32_t i = 0; i < n; i++) {
    const dtwax_ru_chunk& ch = chunks[i];
    auto it = r->chans.find(ch.channel);
    if (it == r->chans.end() || it->second.number != ch.number) {
      if (it != r->chans.end()) release_slot(r, it->second);
      ChanRec fresh;
      fresh.number = ch.number;
      if (r->free_slots.empty()) {
        fresh.action = DTWAX_RU_NO_SLOT;
        r->chans[ch.channel] = fresh;
        set_decision(out[i], ch.channel, fresh, DTWAX_RU_NO_SLOT);
        continue;
      }
      fresh.slot = r->free_slots.back();
      r->free_slots.pop_back();
      r->chans[ch.channel] = fresh;
    } else if (it->second.action != DTWAX_RU_CONTINUE) {
      set_decision(out[i], ch.channel, it->second, DTWAX_RU_STALE);
      continue;
    }
    scored.push_back(i);
  }
  const int64_t m = (int64_t)scored.size();
  if (m == 0) return DTWAX_OK;

  // pass 2: stage samples and work items
  const int64_t N = c->n_cols;
  const int64_t piece = piece_for(r, m);
  const int64_t per_read = (N + piece - 1) / piece;
  if (m * per_read > r->items_cap) return set_error(DTWAX_EINVAL, "work item capacity exceeded");
  int64_t ni = 0;
  for (int64_t j = 0; j < m; j++) {
    const dtwax_ru_chunk& ch = chunks[scored[(size_t)j]];
    const ChanRec& rec = r->chans[ch.channel];
    std::memcpy(r->h_raw + j * L, samples + ch.offset, (size_t)L * sizeof(float));
    r->h_slot[j] = rec.slot;
    const bool first = rec.chunks == 0;
    r->h_first[j] = first ? 1 : 0;
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
      r->h_items[ni++] = w;
    }
  }
  DTWAX_HIP_TRY(hipSetDevice(c->device));
  DTWAX_HIP_TRY(hipMemcpyAsync(r->d_raw, r->h_raw, (size_t)(m * L) * sizeof(float), hipMemcpyHostToDevice, r->st));
  DTWAX_HIP_TRY(hipMemcpyAsync(r->d_slot, r->h_slot, (size_t)m * sizeof(int32_t), hipMemcpyHostToDevice, r->st));
  DTWAX_HIP_TRY(hipMemcpyAsync(r->d_first, r->h_first, (size_t)m * sizeof(int32_t), hipMemcpyHostToDevice, r->st));
  DTWAX_HIP_TRY(hipMemcpyAsync(r->d_items, r->h_items, (size_t)ni * sizeof(WorkItem), hipMemcpyHostToDevice, r->st));
  DTWAX_HIP_TRY(launch_ru_normalize(r->d_raw, (int32_t)m, L, r->d_slot, r->d_first, r->o.norm, r->o.eps,
                                    r->d_stats, r->d_norm, r->st));
  ScoreLaunch sl;
  sl.precision = r->o.precision;
  sl.S = r->S;
  sl.nstr = r->nstr;
  sl.grid = (int)std::min<int64_t>(r->grid_slots, ni);
  sl.queries = r->d_norm;
  sl.q_off = r->d_qoff;
  sl.q_len = r->d_qlen;
  sl.items = r->d_items;
  sl.n_items = (int32_t)ni;
  sl.n_reads = (int32_t)m;
  sl.counter = r->d_counter;
  sl.keys = r->d_keys;
  sl.scratch = r->d_scratch;
  sl.scratch_stride = r->scratch_stride;
  sl.state = (char*)r->d_state + (int64_t)COL_PADL * r->epc * 4;  // column 0 of slot 0, row 0
  sl.row_stride = r->row_stride;
  sl.out = r->d_res;
  sl.ev0 = r->ev0;
  sl.ev1 = r->ev1;
  int rc = run_score_launch(c, sl, r->st);
  if (rc) return rc;
  DTWAX_HIP_TRY(hipMemcpyAsync(r->h_res, r->d_res, (size_t)m * sizeof(DeviceResult), hipMemcpyDeviceToHost, r->st));
  DTWAX_HIP_TRY(hipStreamSynchronize(r->st));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, r->ev0, r->ev1) == hipSuccess) r->gpu_ms += ms;
  r->batches++;
  r->chunks_scored += m;
  r->cells += m * (int64_t)L * N * r->nstr;

  // pass 3: decisions
  for (int64_t j = 0; j < m; j++) {
    const int32_t i = scored[(size_t)j];
    ChanRec& rec = r->chans[chunks[i].channel];
    const DeviceResult& res = r->h_res[j];
    rec.chunks++;
    for (int k = 0; k < 2; k++) {
      rec.score[k] = (k < r->nstr) ? res.score[k] : INFINITY;
      rec.argmin[k] = (k < r->nstr) ? res.argmin[k] : -1;
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
    set_decision(out[i], chunks[i].channel, rec, action);
  }
  return DTWAX_OK;
}

void worker_main(dtwax_ru* r) {
  const int32_t L = r->o.chunk_len;
  std::vector<dtwax_ru_chunk> batch;
  std::vector<float> samples;
  std::vector<Clock::time_point> t_in;
  std::vector<dtwax_ru_decision> dec;
  for (;;) {
    std::vector<Pending> take;
    {
      std::unique_lock<std::mutex> lk(r->mu);
      r->cv_in.wait(lk, [&] { return r->stop_req || !r->inq.empty(); });
      if (r->inq.empty() && r->stop_req) return;
      // every ready chunk, at most one per channel (later ones keep their order)
      std::unordered_map<int32_t, int> seen;
      std::deque<Pending> keep;
      while (!r->inq.empty()) {
        Pending p = std::move(r->inq.front());
        r->inq.pop_front();
        if (seen.emplace(p.channel, 1).second) take.push_back(std::move(p));
        else keep.push_back(std::move(p));
      }
      r->inq.swap(keep);
    }
    const int32_t n = (int32_t)take.size();
    batch.resize((size_t)n);
    samples.resize((size_t)n * L);
    dec.assign((size_t)n, dtwax_ru_decision{});
    for (int32_t i = 0; i < n; i++) {
      batch[(size_t)i].channel = take[(size_t)i].channel;
      batch[(size_t)i].number = take[(size_t)i].number;
      batch[(size_t)i].offset = (int64_t)i * L;
      std::memcpy(samples.data() + (size_t)i * L, take[(size_t)i].samples.data(), (size_t)L * sizeof(float));
    }
    int rc;
    {
      std::lock_guard<std::mutex> lk(r->process_mu);
      rc = process_locked(r, batch.data(), n, samples.data(), dec.data());
    }
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(r->mu);
    if (rc != DTWAX_OK) {
      r->worker_rc = rc;
      r->worker_err = dtwax_last_error();
      r->running = false;
      r->cv_out.notify_all();
      return;
    }
    for (int32_t i = 0; i < n; i++) {
      dec[(size_t)i].latency_us =
          std::chrono::duration<double, std::micro>(now - take[(size_t)i].t).count();
      r->outq.push_back(dec[(size_t)i]);
    }
    r->cv_out.notify_all();
  }
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
  if (o.max_reads < 1 || o.chunk_len < 64 || o.max_chunks < 1)
    return set_error(DTWAX_EINVAL, "max_reads, chunk_len (>= 64) and max_chunks must be positive");
  if (o.precision != DTWAX_F16 && o.precision != DTWAX_F32)
    return set_error(DTWAX_EINVAL, "Read Until serving supports DTWAX_F16 and DTWAX_F32");
  if (o.norm < DTWAX_NORM_NONE || o.norm > DTWAX_NORM_MEDMAD) return set_error(DTWAX_EINVAL, "bad norm");
  if (o.chunk_len % 64) return set_error(DTWAX_EINVAL, "chunk_len must be a multiple of 64");
  const int rows_per_lane = o.chunk_len / 64;
  if (o.segment == 0) {
    for (int S : {32, 25, 16, 8})
      if (rows_per_lane % S == 0) { o.segment = S; break; }
    if (o.segment == 0)
      return set_error(DTWAX_EINVAL, "chunk_len / 64 must be a multiple of 8, 16, 25 or 32");
  }
  if (!valid_segment(o.segment) || rows_per_lane % o.segment)
    return set_error(DTWAX_EINVAL, "segment must be 8, 16, 25 or 32 and divide chunk_len / 64");
  if (o.min_ref_chunk < 0 || o.hbm_budget < 0) return set_error(DTWAX_EINVAL, "negative size");
  if (o.min_ref_chunk == 0) o.min_ref_chunk = 4 * o.chunk_len;
  if (!(o.threshold > 0.0f)) o.threshold = INFINITY;  // no early non-target decision

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
  const int64_t N = c->n_cols;
  // band scratch only for chunks taller than one band: cap the piece there
  r->chunk_cap = r->multiband ? std::min<int64_t>(N, std::max<int64_t>(o.min_ref_chunk, 65536)) : N;
  const int64_t min_piece = std::min<int64_t>(N, std::max<int64_t>(64, (int64_t)o.min_ref_chunk / 64 * 64));
  r->items_cap = slots * ((N + min_piece - 1) / min_piece);
  const int64_t L = o.chunk_len;
  bool ok = hipStreamCreate(&r->st) == hipSuccess && hipEventCreate(&r->ev0) == hipSuccess &&
            hipEventCreate(&r->ev1) == hipSuccess;
  if (!ok) return cleanup(DTWAX_EDEVICE, "stream/event creation failed");
  ok = hipMalloc(&r->d_state, (size_t)(slots * slot_bytes)) == hipSuccess &&
       hipMalloc(&r->d_stats, (size_t)slots * 2 * sizeof(double)) == hipSuccess &&
       hipMalloc(&r->d_raw, (size_t)(slots * L) * sizeof(float)) == hipSuccess &&
       hipMalloc(&r->d_norm, (size_t)(slots * L) * sizeof(float)) == hipSuccess &&
       hipMalloc(&r->d_qoff, (size_t)slots * sizeof(int64_t)) == hipSuccess &&
       hipMalloc(&r->d_qlen, (size_t)slots * sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&r->d_slot, (size_t)slots * sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&r->d_first, (size_t)slots * sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&r->d_items, (size_t)r->items_cap * sizeof(WorkItem)) == hipSuccess &&
       hipMalloc(&r->d_counter, sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&r->d_keys, (size_t)slots * 2 * sizeof(unsigned long long)) == hipSuccess &&
       hipMalloc(&r->d_res, (size_t)slots * sizeof(DeviceResult)) == hipSuccess;
  if (!ok) return cleanup(DTWAX_ENOMEM, "device allocation failed (state " + std::to_string(slots * slot_bytes) + " bytes)");
  ok = hipHostMalloc((void**)&r->h_raw, (size_t)(slots * L) * sizeof(float), 0) == hipSuccess &&
       hipHostMalloc((void**)&r->h_slot, (size_t)slots * sizeof(int32_t), 0) == hipSuccess &&
       hipHostMalloc((void**)&r->h_first, (size_t)slots * sizeof(int32_t), 0) == hipSuccess &&
       hipHostMalloc((void**)&r->h_items, (size_t)r->items_cap * sizeof(WorkItem), 0) == hipSuccess &&
       hipHostMalloc((void**)&r->h_res, (size_t)slots * sizeof(DeviceResult), 0) == hipSuccess;
  if (!ok) return cleanup(DTWAX_ENOMEM, "pinned host allocation failed");
  if (r->multiband) {
    r->scratch_stride = COL_PADL + r->chunk_cap + L + COL_PADR;
    const size_t bytes = (size_t)r->grid_slots * (size_t)r->scratch_stride * (size_t)r->epc * 4;
    if (hipMalloc(&r->d_scratch, bytes) != hipSuccess) return cleanup(DTWAX_ENOMEM, "band scratch allocation failed");
  }
  // state rows start at +inf everywhere: the left pad is column -1 of every
  // stored row (+inf by the recurrence), the rest is overwritten before use
  if (launch_fill_u32((uint32_t*)r->d_state, slots * slot_bytes / 4,
                      o.precision == DTWAX_F32 ? kFloatInf : kHalf2Inf, r->st) != hipSuccess)
    return cleanup(DTWAX_EDEVICE, "state fill failed");
  if (r->d_scratch &&
      launch_fill_u32((uint32_t*)r->d_scratch, r->grid_slots * r->scratch_stride * r->epc, 0u, r->st) != hipSuccess)
    return cleanup(DTWAX_EDEVICE, "scratch fill failed");
  {
    std::vector<int64_t> off((size_t)slots);
    std::vector<int32_t> len((size_t)slots, (int32_t)L);
    for (int64_t j = 0; j < slots; j++) off[(size_t)j] = j * L;
    if (hipMemcpy(r->d_qoff, off.data(), off.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_qlen, len.data(), len.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamSynchronize(r->st) != hipSuccess)
      return cleanup(DTWAX_EDEVICE, "initial upload failed");
  }
  r->free_slots.reserve((size_t)slots);
  for (int64_t s = slots - 1; s >= 0; s--) r->free_slots.push_back((int32_t)s);
  r->slot_cur.assign((size_t)slots, 0);
  *out = r;
  return DTWAX_OK;
}

int dtwax_ru_destroy(dtwax_ru* r) {
  if (!r) return DTWAX_OK;
  (void)dtwax_ru_stop(r);
  ru_free(r);
  return DTWAX_OK;
}

int dtwax_ru_info(const dtwax_ru* r, int32_t* slots, int32_t* active, int32_t* segment, int64_t* state_bytes) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  if (slots) *slots = r->slots;
  if (active) *active = r->slots - (int32_t)r->free_slots.size();
  if (segment) *segment = r->S;
  if (state_bytes) *state_bytes = (int64_t)r->slots * 2 * r->row_stride * r->epc * 4;
  return DTWAX_OK;
}

int dtwax_ru_stats(const dtwax_ru* r, int64_t* batches, int64_t* chunks, int64_t* cells, double* gpu_ms) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  if (batches) *batches = r->batches;
  if (chunks) *chunks = r->chunks_scored;
  if (cells) *cells = r->cells;
  if (gpu_ms) *gpu_ms = r->gpu_ms;
  return DTWAX_OK;
}

int dtwax_ru_process(dtwax_ru* r, const dtwax_ru_chunk* chunks, int32_t n, const float* samples,
                     dtwax_ru_decision* out) {
  if (!r || n < 0 || (n && (!chunks || !samples || !out))) return set_error(DTWAX_EINVAL, "NULL argument");
  for (int32_t i = 0; i < n; i++)
    if (chunks[i].offset < 0 || chunks[i].channel < 0)
      return set_error(DTWAX_EINVAL, "negative channel or offset");
  std::lock_guard<std::mutex> lk(r->process_mu);
  return process_locked(r, chunks, n, samples, out);
}

int dtwax_ru_release(dtwax_ru* r, int32_t channel) {
  if (!r) return set_error(DTWAX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(r->process_mu);
  auto it = r->chans.find(channel);
  if (it != r->chans.end()) {
    release_slot(r, it->second);
    r->chans.erase(it);
  }
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
  p.samples.assign(samples, samples + r->o.chunk_len);
  p.t = Clock::now();
  std::lock_guard<std::mutex> lk(r->mu);
  if (!r->running) return set_error(r->worker_rc ? r->worker_rc : DTWAX_ENOTREADY,
                                    r->worker_rc ? "serving loop failed: " + r->worker_err
                                                 : "serving loop not started");
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
