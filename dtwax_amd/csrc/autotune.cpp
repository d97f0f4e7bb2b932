// autotune.cpp -- picks the lane segment width and the reference chunk of a
// plan by timing the sDTW kernel on a sample of the workload.
//
// The reference has a single compile-time geometry (SEGMENT_SIZE 26, WARP
// 32, one read per block; src/include/common.hpp:60-98) and no tuner.  Here
// the best segment S (rows per lane; a read of Q rows is padded to a multiple
// of 64*S, and the per-step overhead is amortised over S rows) and reference
// chunk (work items per read; each chunk after the first recomputes Q-1
// overlap columns, while too few items starve the persistent grid) depend on
// the length distribution and the reference size.  The width is measured:
// candidate plans run over a sample of reads (spread evenly over the sorted
// lengths, replicated until each candidate has 6 items per resident wave)
// against the first sample_cols reference columns at one moderate chunk, and
// the fastest per counted cell wins.  The chunk is not ranked on the sample:
// a sample holds a few items per wave, so its time is set by how its last
// items land, which favours short chunks the full workload does not need
// (measured on 16,384 reads x 5 Mb: 65536-column chunks won the sample,
// 524288 runs the full workload 6% faster).  It is the plan's own rule at
// the full workload's size with the chosen width: whole strands, or the
// longest chunks whose band scratch fits the budget.  The choice is cached
// per device, precision, strand count, reference-size and length-
// distribution key (in memory, and in a text file when a cache path is
// given).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "dtwax.h"
#include "internal.h"
#include "kernels.h"

using namespace dtwax;

namespace {

std::mutex g_cache_mu;
std::map<std::string, std::pair<int32_t, int32_t>> g_cache;

int64_t bucket_log2(int64_t v) {
  int64_t b = 0;
  while (v > 1) { v >>= 1; b++; }
  return b;
}

std::string tune_key(const dtwax_ctx* c, const dtwax_score_opts& o, const std::vector<int32_t>& sorted_lens) {
  const size_t n = sorted_lens.size();
  auto q = [&](double f) { return sorted_lens[std::min(n - 1, (size_t)(f * (double)(n - 1)))] / 64; };
  std::ostringstream k;
  k << c->name << "/cu" << c->n_cu << "/p" << o.precision << "/S" << o.segment << "/C" << o.ref_chunk << "/m" << (o.moves == 3 ? 3 : 2) << "/k" << o.packing << "/s"
    << c->n_strands << "/N" << bucket_log2(c->n_cols)
    << "/n" << bucket_log2((int64_t)n) << "/q" << q(0.1) << "," << q(0.5) << "," << q(0.9);
  return k.str();
}

bool cache_lookup_file(const char* path, const std::string& key, int32_t& S, int32_t& chunk) {
  if (!path || !*path) return false;
  std::ifstream in(path);
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {  // last entry for a key wins
    std::istringstream ss(line);
    std::string k;
    int32_t s = 0, ch = 0;
    if (ss >> k >> s >> ch && k == key) {
      S = s;
      chunk = ch;
      found = true;
    }
  }
  return found;
}

void cache_store_file(const char* path, const std::string& key, int32_t S, int32_t chunk) {
  if (!path || !*path) return;
  std::ofstream out(path, std::ios::app);
  out << key << " " << S << " " << chunk << "\n";
}

}  // namespace

extern "C" {

int dtwax_autotune(dtwax_ctx* c, const float* d_queries, const int64_t* d_offsets, const int32_t* h_lens,
                   int32_t n_reads, const dtwax_score_opts* base, int32_t sample_reads, int64_t sample_cols,
                   const char* cache_path, dtwax_score_opts* best, dtwax_tune_entry* table,
                   int32_t table_cap, int32_t* n_table) {
  if (!c || !d_queries || !d_offsets || !h_lens || !best || n_reads < 1)
    return set_error(DTWAX_EINVAL, "NULL argument or no reads");
  if (!c->n_cols) return set_error(DTWAX_ENOTREADY, "no reference uploaded");
  if (table_cap < 0 || (table_cap && !table)) return set_error(DTWAX_EINVAL, "bad table");
  dtwax_score_opts o{};
  if (base) o = *base;
  if (o.precision < DTWAX_F16 || o.precision > DTWAX_F32) return set_error(DTWAX_EINVAL, "bad precision");
  if (n_table) *n_table = 0;
  for (int32_t i = 0; i < n_reads; i++)
    if (h_lens[i] < 1) return set_error(DTWAX_EINVAL, "every read needs at least one sample");

  std::vector<int32_t> sorted(h_lens, h_lens + n_reads);
  std::sort(sorted.begin(), sorted.end());
  const std::string key = tune_key(c, o, sorted);
  {
    int32_t S = 0, ch = 0;
    bool hit = false;
    {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      auto it = g_cache.find(key);
      if (it != g_cache.end()) {
        S = it->second.first;
        ch = it->second.second;
        hit = true;
      }
    }
    if (!hit && cache_lookup_file(cache_path, key, S, ch)) {
      hit = true;
      std::lock_guard<std::mutex> lk(g_cache_mu);
      g_cache[key] = {S, ch};
    }
    if (hit) {
      *best = o;
      best->segment = S;
      best->ref_chunk = ch;
      return DTWAX_OK;
    }
  }

  // sample: reads spread evenly over the length order, their offsets copied
  // from the device
  if (sample_reads <= 0) sample_reads = 512;
  const int32_t m = std::min(sample_reads, n_reads);
  std::vector<int32_t> order((size_t)n_reads);
  for (int32_t i = 0; i < n_reads; i++) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return h_lens[a] < h_lens[b]; });
  std::vector<int64_t> all_off((size_t)n_reads);
  DTWAX_HIP_TRY(hipSetDevice(c->device));
  DTWAX_HIP_TRY(hipMemcpy(all_off.data(), d_offsets, (size_t)n_reads * 8, hipMemcpyDeviceToHost));
  std::vector<int64_t> s_off((size_t)m);
  std::vector<int32_t> s_len((size_t)m);
  for (int32_t j = 0; j < m; j++) {
    const int32_t r = order[(size_t)(((int64_t)j * n_reads + n_reads / 2) / m)];
    s_off[(size_t)j] = all_off[(size_t)r];
    s_len[(size_t)j] = h_lens[r];
  }
  if (sample_cols <= 0) sample_cols = (int64_t)1 << 20;
  const int64_t cols = std::min<int64_t>(sample_cols, c->n_cols);

  int64_t* d_soff = nullptr;
  int32_t* d_slen = nullptr;
  dtwax_result* d_res = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (d_soff) (void)hipFree(d_soff);
    if (d_slen) (void)hipFree(d_slen);
    if (d_res) (void)hipFree(d_res);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  if (hipMalloc(&d_soff, (size_t)m * 8) != hipSuccess || hipMalloc(&d_slen, (size_t)m * 4) != hipSuccess ||
      hipMalloc(&d_res, (size_t)m * sizeof(dtwax_result)) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return set_error(DTWAX_ENOMEM, "autotune buffers");
  }
  if (hipMemcpy(d_soff, s_off.data(), (size_t)m * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_slen, s_len.data(), (size_t)m * 4, hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    return set_error(DTWAX_EDEVICE, "autotune upload");
  }

  // 0: per-read widths (the plan's own choice)
  // the renormalised form: its own per-read choice (bands short enough for
  // its precision) or the shortest bands for every read
  std::vector<int32_t> segs = (o.precision == DTWAX_F16_RENORM) ? std::vector<int32_t>{0, 8}
                                                                 : std::vector<int32_t>{0, 8, 12, 16, 20, 25, 28, 32};
  // packing = 1 packs at the widths that have a packed kernel only (automatic, 16..32)
  if (o.packing == 1) segs.erase(std::remove_if(segs.begin(), segs.end(), [](int32_t S) { return S && S < 16; }),
                                 segs.end());
  if (o.segment) segs = {o.segment};
  // Replicated sample: the full workload has far more work items than the
  // sample, so a candidate is timed on enough copies of the sample reads to
  // give it at least kItemsPerWave items per resident wave (steady state,
  // not the tail).
  constexpr int32_t kMaxReps = 64;
  constexpr int64_t kItemsPerWave = 6;
  std::vector<int64_t> r_off((size_t)m * kMaxReps);
  std::vector<int32_t> r_len((size_t)m * kMaxReps);
  for (int32_t k = 0; k < kMaxReps; k++)
    for (int32_t j = 0; j < m; j++) {
      r_off[(size_t)k * m + j] = s_off[(size_t)j];
      r_len[(size_t)k * m + j] = s_len[(size_t)j];
    }
  int64_t* d_roff = nullptr;
  int32_t* d_rlen = nullptr;
  dtwax_result* d_rres = nullptr;
  if (hipMalloc(&d_roff, r_off.size() * 8) != hipSuccess || hipMalloc(&d_rlen, r_len.size() * 4) != hipSuccess ||
      hipMalloc(&d_rres, r_len.size() * sizeof(dtwax_result)) != hipSuccess ||
      hipMemcpy(d_roff, r_off.data(), r_off.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_rlen, r_len.data(), r_len.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    for (void* q : {(void*)d_roff, (void*)d_rlen, (void*)d_rres})
      if (q) (void)hipFree(q);
    cleanup();
    return set_error(DTWAX_ENOMEM, "autotune sample buffers");
  }
  auto free_rep = [&]() {
    for (void* q : {(void*)d_roff, (void*)d_rlen, (void*)d_rres})
      if (q) (void)hipFree(q);
  };

  int32_t nt = 0;
  // times one candidate; returns counted cells per second (< 0: error)
  auto time_candidate = [&](int32_t S, int32_t ch, double& rate) -> int {
    dtwax_score_opts t = o;
    t.segment = S;
    t.ref_chunk = ch;
    dtwax_plan* p = nullptr;
    int rc = plan_build(c, r_len.data(), m, &t, cols, &p);
    if (rc) return rc;
    int64_t grid_sum = 0;
    for (const auto& g : p->groups) grid_sum += g.grid;
    if (p->n_items > 0 && p->n_items < kItemsPerWave * grid_sum) {
      const int32_t reps =
          (int32_t)std::min<int64_t>(kMaxReps, (kItemsPerWave * grid_sum + p->n_items - 1) / p->n_items);
      if (reps > 1) {
        dtwax_plan_destroy(p);
        p = nullptr;
        rc = plan_build(c, r_len.data(), m * reps, &t, cols, &p);
        if (rc) return rc;
      }
    }
    float ms = 0.f;
    rc = dtwax_score_plan(p, d_queries, d_roff, d_rlen, d_rres, st);  // warm-up
    if (!rc) {
      hipError_t e = hipEventRecord(e0, st);
      if (e == hipSuccess) rc = dtwax_score_plan(p, d_queries, d_roff, d_rlen, d_rres, st);
      if (!rc && e == hipSuccess) e = hipEventRecord(e1, st);
      if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
      if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (!rc && e != hipSuccess) rc = set_error(DTWAX_EDEVICE, std::string("autotune: ") + hipGetErrorString(e));
    }
    rate = ms > 0 ? (double)p->cells / (ms * 1e-3) : 0.0;
    dtwax_plan_destroy(p);
    if (rc) return rc;
    if (nt < table_cap) {
      table[nt].segment = S;
      table[nt].ref_chunk = ch;
      table[nt].ms = ms;
      table[nt].gcups = (float)(rate / 1e9);
    }
    nt++;
    return DTWAX_OK;
  };

  // widths, timed at one moderate chunk (the three-move recurrence runs whole
  // strands: its left move crosses chunks)
  const bool m3 = o.moves == 3;
  const int32_t chunk0 = m3 ? 0 : o.ref_chunk ? o.ref_chunk : (int32_t)std::min<int64_t>(65536, cols);
  // Segment 0 (each read at its best width, packed partial bands) is the
  // plan's own optimum under its cost model, so a fixed width replaces it
  // only when it measures kFixedMargin faster on the sample: on the uneven
  // 16,384-read 5 Mb workload a fixed S = 8 led the sample by 0.9% and ran
  // the full workload 6% slower.
  constexpr double kFixedMargin = 1.02;
  double best_rate = -1.0, rate0 = -1.0;
  int32_t best_S = segs[0], best_chunk = chunk0;
  for (int32_t S : segs) {
    double rate = 0;
    int rc = time_candidate(S, chunk0, rate);
    if (rc) {
      free_rep();
      cleanup();
      return rc;
    }
    if (S == 0) rate0 = rate;
    if (rate > best_rate) {
      best_rate = rate;
      best_S = S;
    }
  }
  if (rate0 > 0 && best_S != 0 && best_rate < kFixedMargin * rate0) best_S = 0;
  free_rep();
  cleanup();
  if (!o.ref_chunk && !m3) {
    // the plan's chunk rule on the full workload with the chosen width
    dtwax_score_opts t = o;
    t.segment = best_S;
    t.ref_chunk = 0;
    dtwax_plan* p = nullptr;
    int rc = plan_build(c, h_lens, n_reads, &t, c->n_cols, &p, /*dry=*/true);
    if (rc) return rc;
    best_chunk = p->chunk;
    dtwax_plan_destroy(p);
  }
  if (n_table) *n_table = std::min(nt, table_cap);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache[key] = {best_S, best_chunk};
  }
  cache_store_file(cache_path, key, best_S, best_chunk);
  *best = o;
  best->segment = best_S;
  best->ref_chunk = best_chunk;
  return DTWAX_OK;
}

}  // extern "C"
