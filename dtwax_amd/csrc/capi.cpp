// capi.cpp -- host side of the C ABI declared in include/dtwax.h.
//
// Owns the device copies of the reference, plans the (read, column-chunk)
// work items for the persistent sDTW kernel, and maps every HIP error to a
// status code with a message (the reference instead exits on the first error
// it sees, src/include/hpc_helpers.hpp:103-113).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dtwax.h"
#include "internal.h"
#include "kernels.h"

using namespace dtwax;

namespace {
thread_local std::string g_err;
inline hipStream_t S_(void* s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace

namespace dtwax {
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace dtwax

namespace {
int fail(int code, const std::string& msg) { return set_error(code, msg); }
}  // namespace

#define HIP_TRY(expr) DTWAX_HIP_TRY(expr)

// ---- scratch of the allocation-free entry point -----------------------------
// dtwax_distances_noalloc keeps everything beyond its inputs and outputs in
// this file-scope device array, so no call allocates: the untiled, padded
// reference (FP16 build: 64 + 47,424 + 384 half2 words), one work counter
// and 2 keys per read for at most kNoAllocMaxEntries reads, 448 KB per GPU
// for the life of the process.  Host code never uses the name `scratch`
// itself (on the host it is only a stand-in for the symbol): each device's
// address comes from hipGetSymbolAddress once, in dtwax_create with that
// device current, and is kept per device.
namespace {
constexpr int32_t kNoAllocMaxEntries = 16384;
constexpr size_t kNoAllocRefWords = COL_PADL + 47424 + COL_PADR;
constexpr size_t kNoAllocRefBytes = (kNoAllocRefWords * 4 + 255) / 256 * 256;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr size_t kNoAllocKeyBytes = (size_t)kNoAllocMaxEntries * 2 * sizeof(unsigned long long);
constexpr size_t kNoAllocBytes = kNoAllocRefBytes + kNoAllocCounterBytes + kNoAllocKeyBytes;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];


extern "C" {

const char* dtwax_last_error(void) { return g_err.c_str(); }
const char* dtwax_version(void) { return "dtwax-mi355x 0.1 (gfx950)"; }

int dtwax_create(int device, dtwax_ctx** out) {
  if (!out) return fail(DTWAX_EINVAL, "out is NULL");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(DTWAX_EINVAL, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* ptr = nullptr;
      HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      if (!ptr) return fail(DTWAX_EDEVICE, "hipGetSymbolAddress(scratch) gave no address");
      g_scratch_ptr[device] = ptr;
    }
  }
  auto* c = new dtwax_ctx;
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  c->clock_khz = prop.clockRate;
  c->name = prop.gcnArchName;
  *out = c;
  return DTWAX_OK;
}

static void free_reference(dtwax_ctx* c) {
  if (c->ref16_alloc) (void)hipFree(c->ref16_alloc);
  if (c->ref32_alloc) (void)hipFree(c->ref32_alloc);
  c->ref16_alloc = c->ref32_alloc = nullptr;
  c->n_cols = 0;
  c->n_strands = 0;
}

void dtwax_destroy(dtwax_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  free_reference(c);
  for (void* p : {c->compat_ref[0], c->compat_ref[1], (void*)c->compat_counter, (void*)c->compat_keys})
    if (p) (void)hipFree(p);
  delete c;
}

int dtwax_device_info(dtwax_ctx* c, int32_t* n_cu, int32_t* clock_khz, char* name, int32_t name_len) {
  if (!c) return fail(DTWAX_EINVAL, "ctx is NULL");
  if (n_cu) *n_cu = c->n_cu;
  if (clock_khz) *clock_khz = c->clock_khz;
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s", c->name.c_str());
  }
  return DTWAX_OK;
}

// ---- memory / streams / events -------------------------------------------
int dtwax_malloc(void** p, size_t bytes) {
  if (!p) return fail(DTWAX_EINVAL, "ptr is NULL");
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(DTWAX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return DTWAX_OK;
}
int dtwax_free(void* p) {
  HIP_TRY(hipFree(p));
  return DTWAX_OK;
}
int dtwax_memcpy_h2d(void* dst, const void* src, size_t bytes, void* st) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S_(st)));
  HIP_TRY(hipStreamSynchronize(S_(st)));
  return DTWAX_OK;
}
int dtwax_memcpy_d2h(void* dst, const void* src, size_t bytes, void* st) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S_(st)));
  HIP_TRY(hipStreamSynchronize(S_(st)));
  return DTWAX_OK;
}
int dtwax_stream_create(void** st) {
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *st = s;
  return DTWAX_OK;
}
int dtwax_stream_destroy(void* st) {
  HIP_TRY(hipStreamDestroy(S_(st)));
  return DTWAX_OK;
}
int dtwax_stream_sync(void* st) {
  HIP_TRY(hipStreamSynchronize(S_(st)));
  return DTWAX_OK;
}
int dtwax_device_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return DTWAX_OK;
}
int dtwax_event_create(void** ev) {
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  *ev = e;
  return DTWAX_OK;
}
int dtwax_event_destroy(void* ev) {
  HIP_TRY(hipEventDestroy((hipEvent_t)ev));
  return DTWAX_OK;
}
int dtwax_event_record(void* ev, void* st) {
  HIP_TRY(hipEventRecord((hipEvent_t)ev, S_(st)));
  return DTWAX_OK;
}
int dtwax_event_elapsed_ms(void* a, void* b, float* ms) {
  HIP_TRY(hipEventSynchronize((hipEvent_t)b));
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return DTWAX_OK;
}

// ---- reference -----------------------------------------------------------
static const _Float16 kPadF16 = (_Float16)65504.0f;  // (pad - q)^2 overflows to +inf
static const float kPadF32 = 1e30f;                   // (pad - q)^2 overflows to +inf

int dtwax_reference_upload(dtwax_ctx* c, const float* levels, int32_t n_strands, int64_t n_cols) {
  if (!c || !levels) return fail(DTWAX_EINVAL, "NULL argument");
  if (n_strands < 1 || n_strands > 2) return fail(DTWAX_EINVAL, "n_strands must be 1 or 2");
  if (n_cols < 1 || n_cols > (int64_t)INT32_MAX - 4096) return fail(DTWAX_EINVAL, "n_cols out of range");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipDeviceSynchronize());
  free_reference(c);
  const int64_t total = COL_PADL + n_cols + COL_PADR;
  std::vector<uint32_t> h16((size_t)total);
  std::vector<float> h32((size_t)(total * n_strands));
  for (int64_t i = 0; i < total; i++) {
    const int64_t col = i - COL_PADL;
    const bool real = col >= 0 && col < n_cols;
    _Float16 a = real ? (_Float16)levels[col] : kPadF16;
    _Float16 b = (real && n_strands == 2) ? (_Float16)levels[n_cols + col] : kPadF16;
    h16[(size_t)i] = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    for (int k = 0; k < n_strands; k++)
      h32[(size_t)(i * n_strands + k)] = real ? levels[k * n_cols + col] : kPadF32;
  }
  hipError_t e1 = hipMalloc(&c->ref16_alloc, (size_t)total * 4);
  hipError_t e2 = hipMalloc(&c->ref32_alloc, (size_t)total * 4 * n_strands);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    free_reference(c);
    return fail(DTWAX_ENOMEM, "reference allocation failed");
  }
  HIP_TRY(hipMemcpy(c->ref16_alloc, h16.data(), (size_t)total * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->ref32_alloc, h32.data(), (size_t)total * 4 * n_strands, hipMemcpyHostToDevice));
  c->n_cols = n_cols;
  c->n_strands = n_strands;
  return DTWAX_OK;
}

// ---- normalisation -------------------------------------------------------
int dtwax_normalize(dtwax_ctx* c, const float* d_raw, const int64_t* d_off, const int32_t* d_len,
                    int32_t n_reads, int32_t mode, double eps, float* d_out, double* d_stats, void* st) {
  if (!c || (!d_raw && n_reads) || (!d_out && n_reads)) return fail(DTWAX_EINVAL, "NULL argument");
  if (mode < DTWAX_NORM_NONE || mode > DTWAX_NORM_MEDMAD) return fail(DTWAX_EINVAL, "bad mode");
  if (n_reads < 0) return fail(DTWAX_EINVAL, "n_reads < 0");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_normalize(d_raw, d_off, d_len, n_reads, mode, eps, d_out, d_stats, S_(st)));
  return DTWAX_OK;
}

// ---- planning ------------------------------------------------------------
}  // extern "C"

int dtwax::occupancy(int precision, int S, int nstr, bool m3) {
  return precision == DTWAX_F32 ? occupancy_f32(S, nstr, m3) : occupancy_f16(precision, S, m3);
}

extern "C" {

// Bytes of band scratch a plan may take: a quarter of the free device
// memory, at most 32 GB (on a 288 GB device: 512k-column chunks for a full
// grid of 8000-sample reads, whose Q - 1 overlap re-computes < 2% of the
// cells), at least 1 GB.
static int64_t scratch_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)32 << 30;
  return std::max<int64_t>((int64_t)1 << 30, std::min<int64_t>((int64_t)32 << 30, (int64_t)(free_b / 4)));
}

static bool valid_segment(int S) {
  return S == 8 || S == 12 || S == 16 || S == 20 || S == 25 || S == 28 || S == 32;
}

static int auto_segment(const std::vector<int32_t>& lens) {
  // binary32: one width per batch, the least padded rows x the measured time
  // per computed cell of that width relative to S = 16 (4096 x 2000-sample
  // batches, band factor taken out: 1.113 at 8, 1.060 at 12, 1.097 at 20,
  // 0.996 at 25, 0.984 at 28, 1.006 at 32; from S = 20 up 2 waves fit a SIMD,
  // at S = 16 three, and from S = 25 up the row costs are looked ahead across
  // steps) x 1.5% per extra band
  static const int cand[] = {8, 12, 16, 20, 25, 28, 32};
  static const double f0[] = {1.113, 1.060, 1.0, 1.097, 0.996, 0.984, 1.006};
  double best = 1e300;
  int bs = 16;
  for (int ci = 0; ci < 7; ci++) {
    const int64_t BR = 64 * cand[ci];
    double cost = 0;
    for (int32_t q : lens) {
      const int64_t nb = ((int64_t)q + BR - 1) / BR;
      cost += (double)(nb * BR) * f0[ci] * (1.0 + 0.015 * (double)(nb - 1));
    }
    if (cost < best - 1e-9) {
      best = cost;
      bs = cand[ci];
    }
  }
  return bs;
}

int dtwax_plan_create(dtwax_ctx* c, const int32_t* h_lens, int32_t n_reads, const dtwax_score_opts* opts,
                      dtwax_plan** out) {
  return dtwax::plan_build(c, h_lens, n_reads, opts, c ? c->n_cols : 0, out);
}

}  // extern "C"

namespace {
// Binary16 segment widths and each one's time per computed cell relative to
// S = 32, measured on 4096 x 2000-sample batches with the band factor
// taken out.
constexpr int kWidths[7] = {8, 12, 16, 20, 25, 28, 32};
constexpr double kWidthCost[7] = {1.03, 1.011, 1.011, 1.0, 1.0, 1.0, 1.0};
// The packed layout's shared band: 1.6% more packed ops per cell (lane
// masks beside the DPP moves) and the history flushes.
constexpr double kPackCost = 1.02;

// Padded work of a q-row read in bands of 64 x S rows (S one of kWidths):
// rows x the width's cell time x 1.5% per extra band (hand-over through
// scratch).
double band_cost(int64_t q, int S) {
  const int ci = (int)(std::find(kWidths, kWidths + 7, S) - kWidths);
  const int64_t BR = 64 * (int64_t)S;
  const int64_t nb = (q + BR - 1) / BR;
  return (double)(nb * BR) * kWidthCost[ci] * (1.0 + 0.015 * (double)(nb - 1));
}

// The width with the least padded work (the narrower one on ties).
int best_width(int64_t q) {
  int best = kWidths[0];
  for (int ci = 1; ci < 7; ci++)
    if (band_cost(q, kWidths[ci]) < band_cost(q, best) - 1e-9) best = kWidths[ci];
  return best;
}

// A packed read: `rp` rows (its length modulo 64 x SP, or a whole band) in
// `lanes` lanes of a shared band, then `nfull` full bands.
struct PackSeg {
  int32_t read, lanes, nfull, rp;
};
struct PackBin {
  int32_t used = 0, nmulti = 0;  // lanes taken; reads continuing past the band
  std::vector<int32_t> segs;     // indices into the segment list, lane order
};

// Picks the reads to pack and places their partial bands in 64-lane bands.
// A read needs at least 4 lanes (the history flush period is at least 4
// steps).  Automatic: the reads whose packed cost beats their unpacked width
// by 1%, kept only if the bands as filled still beat those reads' unpacked
// cost.  Placement: best fit decreasing, at most max_multi continuing reads
// per band (each needs a band scratch row per wave).  Returns the packed
// reads' cost (band_cost units), 0 when nothing is packed.
double plan_packing(const std::vector<int32_t>& lens, const std::vector<int>& seg_of, int SP, bool automatic,
                    int max_multi, std::vector<PackSeg>& segs, std::vector<PackBin>& bins) {
  const int64_t BRP = 64 * (int64_t)SP;
  for (int32_t r = 0; r < (int32_t)lens.size(); r++) {
    const int64_t q = lens[(size_t)r];
    const int32_t nfull = (int32_t)((q - 1) / BRP);
    const int32_t rp = (int32_t)(q - nfull * BRP);
    const int32_t lanes = (rp + SP - 1) / SP;
    if (lanes < 4) continue;
    const double pk = ((double)(nfull * BRP) + kPackCost * lanes * SP) * (1.0 + 0.015 * nfull);
    if (automatic && pk >= 0.99 * band_cost(q, seg_of[(size_t)r])) continue;
    segs.push_back({r, lanes, nfull, rp});
  }
  if (segs.empty()) return 0;
  std::stable_sort(segs.begin(), segs.end(), [](const PackSeg& a, const PackSeg& b) { return a.lanes > b.lanes; });
  std::set<std::pair<int32_t, int32_t>> open_any, open_final;  // (free lanes, bin)
  for (int32_t i = 0; i < (int32_t)segs.size(); i++) {
    const PackSeg& sg = segs[(size_t)i];
    const bool multi = sg.nfull > 0;
    int32_t b = -1;
    const auto ia = open_any.lower_bound({sg.lanes, -1});
    const auto jf = multi ? open_final.end() : open_final.lower_bound({sg.lanes, -1});  // full of continuing reads
    const bool has_a = ia != open_any.end(), has_f = jf != open_final.end();
    if (has_f && (!has_a || jf->first < ia->first)) {
      b = jf->second;
      open_final.erase(jf);
    } else if (has_a) {
      b = ia->second;
      open_any.erase(ia);
    } else {
      b = (int32_t)bins.size();
      bins.emplace_back();
    }
    PackBin& bn = bins[(size_t)b];
    bn.used += sg.lanes;
    bn.nmulti += multi ? 1 : 0;
    bn.segs.push_back(i);
    if (bn.used < 64) (bn.nmulti < max_multi ? open_any : open_final).insert({64 - bn.used, b});
  }
  double packed = kPackCost * (double)BRP * (double)bins.size(), unpacked = 0;
  for (const PackSeg& sg : segs) {
    unpacked += band_cost(lens[(size_t)sg.read], seg_of[(size_t)sg.read]);
    packed += (double)(sg.nfull * BRP) * (1.0 + 0.015 * sg.nfull);
  }
  if (automatic && packed >= 0.99 * unpacked) {
    segs.clear();
    bins.clear();
    return 0;
  }
  return packed;
}
}  // namespace

int dtwax::plan_build(dtwax_ctx* c, const int32_t* h_lens, int32_t n_reads, const dtwax_score_opts* opts,
                      int64_t col_limit, dtwax_plan** out, bool dry) {
  if (!c || !out || (!h_lens && n_reads)) return fail(DTWAX_EINVAL, "NULL argument");
  if (col_limit < 1 || col_limit > c->n_cols) return fail(DTWAX_EINVAL, "bad column limit");
  if (n_reads < 0) return fail(DTWAX_EINVAL, "n_reads < 0");
  if (!c->n_cols) return fail(DTWAX_ENOTREADY, "no reference uploaded");
  dtwax_score_opts o{};
  if (opts) o = *opts;
  if (o.precision < DTWAX_F16 || o.precision > DTWAX_F32) return fail(DTWAX_EINVAL, "bad precision");
  if (o.segment && !valid_segment(o.segment)) return fail(DTWAX_EINVAL, "segment must be 8, 12, 16, 20, 25, 28 or 32");
  if (o.ref_chunk < 0) return fail(DTWAX_EINVAL, "ref_chunk < 0");
  if (o.waves < 0) return fail(DTWAX_EINVAL, "waves < 0");
  if (o.compat_quirks) return fail(DTWAX_EINVAL, "compat_quirks is only available through dtwax_distances");
  if (o.moves != 0 && o.moves != 2 && o.moves != 3) return fail(DTWAX_EINVAL, "moves must be 2 or 3");
  const bool m3 = o.moves == 3;
  if (m3 && o.precision != DTWAX_F16 && o.precision != DTWAX_F32)
    return fail(DTWAX_EINVAL, "moves = 3 needs DTWAX_F16 or DTWAX_F32");
  if (m3 && o.ref_chunk && o.ref_chunk < col_limit)
    return fail(DTWAX_EINVAL, "moves = 3 needs whole-strand work items (ref_chunk 0)");
  std::vector<int32_t> lens(h_lens, h_lens + n_reads);
  int32_t qmax = 0;
  for (int32_t q : lens) {
    if (q < 1) return fail(DTWAX_EINVAL, "every read needs at least one sample");
    qmax = std::max(qmax, q);
  }
  if (o.packing < -1 || o.packing > 1) return fail(DTWAX_EINVAL, "packing must be -1, 0 or 1");
  const bool pack_form = !m3 && (o.precision == DTWAX_F16 || o.precision == DTWAX_F16_FUSED);
  if (o.packing == 1 && (!pack_form || (o.segment && o.segment < 16)))
    return fail(DTWAX_EINVAL, "packing needs DTWAX_F16 or DTWAX_F16_FUSED, two moves and segment 0 or 16..32");
  HIP_TRY(hipSetDevice(c->device));

  auto* p = new dtwax_plan;
  p->ctx = c;
  p->n_reads = n_reads;
  p->precision = o.precision;
  p->nstr = c->n_strands;
  p->m3 = m3;
  // initial per-row drift; each band then adds its exact row minimum / 64 per
  // lane.  Measured against FP64 over Q = 1000..8000 (both classes): 0.01
  // keeps the max relative error below 1%; 0.1 over-subtracts for reads
  // whose row minimum grows slowly (targets) and loses up to 15% at Q = 1000.
  const float drift = o.renorm_drift > 0 ? o.renorm_drift : 0.01f;
  const int64_t N = col_limit;

  // Segment width per read.  Fixed when given; binary32 takes one width for
  // the batch; the binary16 forms pick, per read, the width with the least
  // padded work.  The renormalised form subtracts each band's exact row
  // minimum from the next band, so its first band must be short enough that
  // binary16 has not drifted yet: reads of up to 1536 rows take bands of
  // 64 x 8 rows (two or more bands); longer ones run the S = 16 kernel's
  // split schedule (their first 513..1536 rows in bands of 64 x 8 rows, the
  // rest in bands of 64 x 16 rows on the two-column sweep).  Against FP64
  // (max relative error of min(fwd, rev), 1024 reads of the 48.5 kb
  // workload): 0.75% / 0.58% / 0.55% / 0.72% at 1000 / 2000 / 4000 / 8000
  // rows; a first band of 1024 rows instead gave 1.8% at 2000 rows.
  std::vector<int> seg_of((size_t)n_reads, 0);
  if (o.segment) {
    std::fill(seg_of.begin(), seg_of.end(), o.segment);
  } else if (o.precision == DTWAX_F16_RENORM) {
    for (int32_t r = 0; r < n_reads; r++) seg_of[(size_t)r] = lens[r] <= 1536 ? 8 : 16;
  } else if (o.precision == DTWAX_F32) {
    std::fill(seg_of.begin(), seg_of.end(), auto_segment(lens));
  } else {
    for (int32_t r = 0; r < n_reads; r++) seg_of[(size_t)r] = best_width(lens[r]);
  }
  // Packing width: the segment when given, else the width whose packing
  // (with the other reads at their own widths) costs least.  Up to 4
  // continuing reads per band when whole-strand rows for them fit the
  // scratch budget, else 2 (a chunk shortened for the rows costs more
  // overlap than the extra reads save).
  int SP = o.segment ? o.segment : 32;
  std::vector<PackSeg> segs;
  std::vector<PackBin> bins;
  if (pack_form && (o.packing == 1 || (o.packing == 0 && !o.segment))) {
    auto max_multi = [&](int S) {
      const int64_t slots_p = (int64_t)std::max(1, occupancy_f16_pack(o.precision, S)) * c->n_cu;
      const int64_t row = (COL_PADL + N + qmax + COL_PADR) * 4;
      return row * 4 * slots_p <= scratch_budget() ? 4 : 2;
    };
    if (o.packing == 1) {
      plan_packing(lens, seg_of, SP, false, max_multi(SP), segs, bins);
    } else {
      double best = 0;
      for (int32_t r = 0; r < n_reads; r++) best += band_cost(lens[r], seg_of[(size_t)r]);
      for (int S : {32, 28, 25, 20, 16}) {
        std::vector<PackSeg> sg;
        std::vector<PackBin> bn;
        double cost = plan_packing(lens, seg_of, S, true, max_multi(S), sg, bn);
        if (sg.empty()) continue;
        std::vector<char> in((size_t)n_reads, 0);
        for (const PackSeg& x : sg) in[(size_t)x.read] = 1;
        for (int32_t r = 0; r < n_reads; r++)
          if (!in[(size_t)r]) cost += band_cost(lens[r], seg_of[(size_t)r]);
        if (cost < 0.995 * best) {
          best = cost;
          SP = S;
          segs.swap(sg);
          bins.swap(bn);
        }
      }
    }
  }
  const int64_t BRP = 64 * SP;
  std::vector<char> packed_read((size_t)n_reads, 0);
  for (const PackSeg& sg : segs) packed_read[(size_t)sg.read] = 1;

  std::vector<int> widths;
  for (int32_t r = 0; r < n_reads; r++)
    if (!packed_read[(size_t)r] && std::find(widths.begin(), widths.end(), seg_of[(size_t)r]) == widths.end())
      widths.push_back(seg_of[(size_t)r]);
  std::sort(widths.rbegin(), widths.rend());
  if (widths.empty() && bins.empty()) widths.push_back(o.segment ? o.segment : 32);
  if (bins.empty()) {
    p->S = widths.size() == 1 ? widths[0] : 0;
  } else {
    p->S = (widths.empty() || (widths.size() == 1 && widths[0] == SP)) ? SP : 0;
  }
  p->renorm_k = drift * (float)(widths.empty() ? SP : widths[0]);

  bool multiband = false;
  int32_t pack_rows = 1;  // band scratch rows per workgroup of the packed launch
  for (const PackBin& bn : bins) pack_rows = std::max(pack_rows, bn.nmulti);
  int64_t slots = 1, slots_sum = 0;
  for (int32_t r = 0; r < n_reads; r++) multiband |= !packed_read[(size_t)r] && lens[r] > 64 * seg_of[(size_t)r];
  for (const PackSeg& sg : segs) multiband |= sg.nfull > 0;
  for (int S : widths) {
    const int64_t sl = (int64_t)std::max(1, dtwax::occupancy(p->precision, S, p->nstr, m3)) * c->n_cu;
    slots = std::max(slots, sl);
    slots_sum += sl;
  }
  const int64_t slots_pack = bins.empty() ? 0 : (int64_t)std::max(1, occupancy_f16_pack(p->precision, SP)) * c->n_cu;
  if (!bins.empty()) {
    slots = std::max(slots, slots_pack);
    slots_sum += slots_pack * pack_rows;
  }
  const int64_t n_units = (int64_t)(n_reads - (int32_t)segs.size()) + (int64_t)bins.size();

  // chunking: explicit, or whole strands unless the GPU would be underfilled
  // or the band scratch would not fit the budget.
  int64_t chunk = o.ref_chunk ? o.ref_chunk : N;
  if (!o.ref_chunk && !m3) {
    if (n_units < slots && n_units > 0 && p->precision == DTWAX_F16_RENORM) {
      // the renormalised form's rounding depends on its chunks (each band
      // subtracts its own item's row minimum): its automatic plans keep
      // this rule
      int64_t want = (slots + n_units - 1) / n_units;
      int64_t ch = (N + want - 1) / want;
      chunk = std::max<int64_t>(ch, std::min<int64_t>(N, 8 * (int64_t)qmax));
    } else if (n_units < slots && n_units > 0) {
      // K equal pieces per read, K by the modelled time of the slowest SIMD:
      // items per SIMD x (piece + Q - 1) / SIMD throughput at that many
      // waves.  One binary16 sDTW wave alone on a SIMD does nearly the work
      // of three sharing it (48.5 kb x 2000-sample reads at S = 32: 1.0 /
      // 1.05 / 1.06), so a small batch gains from long pieces, one per SIMD,
      // and loses to pieces that put a second item on some SIMDs.  Measured
      // against max(N / (slots / n), 8 Q) chunks: 256 reads 4.42 -> 3.47 ms,
      // 512 reads 8.47 -> 6.39, 1024 reads 12.48 -> 11.69, 2048 reads
      // 23.42 -> 22.33, 1536 and 2560 reads unchanged.  Binary32 ops issue
      // side by side only from two waves of a SIMD (a lone wave: about 0.4
      // of the pair's rate), so binary32 counts 2 for two or more waves.
      static const double kRateF16[4] = {0.0, 1.0, 1.05, 1.06};
      static const double kRateF32[4] = {0.0, 1.0, 2.0, 2.0};
      const double* kRate = p->precision == DTWAX_F32 ? kRateF32 : kRateF16;
      const int64_t n_simd = 4 * (int64_t)c->n_cu;
      double t_best = 0.0;
      int64_t k_best = 1;
      for (int64_t K = 1; K <= N && n_units * K <= 3 * slots; K++) {
        const int64_t piece = (N + K - 1) / K;
        if (K > 1 && piece < 64) break;
        const int64_t items = n_units * K;
        const int64_t per_simd = (items + n_simd - 1) / n_simd;
        const int64_t w = std::min<int64_t>(3, (std::min<int64_t>(items, slots) + n_simd - 1) / n_simd);
        const int64_t cols = K == 1 ? N : piece + qmax - 1;
        const double t = (double)per_simd * (double)cols / kRate[w];
        if (K == 1 || t < t_best) {
          t_best = t;
          k_best = K;
        }
      }
      chunk = (N + k_best - 1) / k_best;
    }
    if (multiband) {
      const int64_t budget = scratch_budget();
      const int64_t bytes_per_col = 4 * (p->precision == DTWAX_F32 ? p->nstr : 1);
      const int64_t max_w = budget / (std::min<int64_t>(slots_sum, n_units * pack_rows * ((N + chunk - 1) / chunk)) * bytes_per_col);
      if (chunk + qmax > max_w) chunk = std::max<int64_t>(4096, max_w - qmax);
    }
  }
  chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, N));
  p->chunk = (int32_t)chunk;

  std::vector<WorkItem> hv;
  int64_t max_w = 0, scratch_rows = 0;
  int max_grid = 1;
  const int strands_counted = c->n_strands;
  for (int S : widths) {
    const int64_t BR = 64 * S;
    std::vector<std::pair<int64_t, WorkItem>> items;
    for (int32_t r = 0; r < n_reads; r++) {
      if (seg_of[(size_t)r] != S || packed_read[(size_t)r]) continue;
      const int64_t q = lens[r];
      // rows computed, virtual ones included (the split schedule's first
      // part pads to 512 rows)
      int64_t rows = (q + BR - 1) / BR * BR;
      if (o.precision == DTWAX_F16_RENORM && !o.segment && S == 16 && q > 1536) {
        const int64_t rest = 1024 * ((q - 513) / 1024);
        rows = rest + (q - rest + 511) / 512 * 512;
      }
      for (int64_t c0 = 0; c0 < N; c0 += chunk) {
        const int64_t c1 = std::min<int64_t>(N, c0 + chunk);
        const int64_t x = std::max<int64_t>(0, c0 - (q - 1));
        WorkItem it{};
        it.read = r;
        it.slot = 0;
        it.x = (int32_t)x;
        it.c0 = (int32_t)c0;
        it.c1 = (int32_t)c1;
        const int64_t steps = ((c1 - x + 63) + 63) / 64 * 64;
        const int64_t cost = rows * steps;
        items.push_back({cost, it});
        p->cells += q * (c1 - c0) * strands_counted;
        p->computed += rows * steps * (p->precision == DTWAX_F32 ? p->nstr : 2);
        max_w = std::max(max_w, c1 - x);
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<int64_t, WorkItem>& a, const std::pair<int64_t, WorkItem>& b) {
                       return a.first > b.first;
                     });
    LaunchGroup g;
    g.S = S;
    g.item0 = (int32_t)hv.size();
    g.n_items = (int32_t)items.size();
    const int64_t slots_S = (int64_t)std::max(1, dtwax::occupancy(p->precision, S, p->nstr, m3)) * c->n_cu;
    g.grid = (int)std::max<int64_t>(1, std::min<int64_t>(slots_S, (int64_t)items.size()));
    if (o.waves > 0) g.grid = std::min(g.grid, o.waves);
    g.renorm_k = drift * (float)S;
    if (o.precision == DTWAX_F16_RENORM && !o.segment && S == 16) {
      g.renorm_split = true;  // starts at 64 x 8 rows per band
      g.renorm_k = drift * 8.0f;
    }
    g.scratch_wg0 = scratch_rows;  // concurrent groups: disjoint band scratch
    scratch_rows += g.grid;
    max_grid = std::max(max_grid, g.grid);
    for (auto& pr : items) hv.push_back(pr.second);
    p->groups.push_back(g);
  }
  // packed items: one per band of packed partial bands and chunk
  std::vector<PackLane> hlanes;
  std::vector<PackRead> hreads;
  if (!bins.empty()) {
    std::vector<std::pair<int64_t, WorkItem>> items;
    for (int32_t b = 0; b < (int32_t)bins.size(); b++) {
      const PackBin& bn = bins[(size_t)b];
      const int32_t first_read = (int32_t)hreads.size();
      int32_t lane = 0, min_lanes = 64, nmulti = 0;
      int64_t qb = 1;
      for (int32_t si : bn.segs) {
        const PackSeg& sg = segs[(size_t)si];
        const int32_t P = sg.lanes * SP - sg.rp;
        const int32_t endl = lane + sg.lanes - 1;
        for (int32_t l = 0; l < sg.lanes; l++) {
          PackLane pl;
          pl.read = sg.read;
          pl.r0 = l * SP - P;
          pl.info = endl | (l == 0 ? PL_START : 0) | (sg.nfull ? PL_HAND : PL_FINAL);
          pl.row = sg.nfull ? nmulti : -1;
          hlanes.push_back(pl);
        }
        if (sg.nfull) {
          hreads.push_back({sg.read, sg.nfull, sg.rp, nmulti});
          nmulti++;
        }
        lane += sg.lanes;
        min_lanes = std::min(min_lanes, sg.lanes);
        qb = std::max<int64_t>(qb, lens[sg.read]);
      }
      for (; lane < 64; lane++) hlanes.push_back({-1, -SP, lane, -1});  // idle
      std::vector<std::pair<int64_t, int32_t>> multi;  // (length, full bands) of continuing reads
      for (int32_t si : bn.segs)
        if (segs[(size_t)si].nfull) multi.push_back({lens[segs[(size_t)si].read], segs[(size_t)si].nfull});
      int32_t F = 64;
      while (F > min_lanes) F >>= 1;
      for (int64_t c0 = 0; c0 < N; c0 += chunk) {
        const int64_t c1 = std::min<int64_t>(N, c0 + chunk);
        const int64_t x = std::max<int64_t>(0, c0 - (qb - 1));
        WorkItem it{};
        it.read = b;
        it.slot = first_read;
        it.x = (int32_t)x;
        it.c0 = (int32_t)c0;
        it.c1 = (int32_t)c1;
        it.flags = ITEM_PACK;
        it.cur = nmulti;
        it.reserved = F;
        const int64_t steps = ((c1 - x + 63) + 63) / 64 * 64;
        // the shared band from x; each read's full bands from its own overlap
        int64_t band_steps = steps;
        for (const auto& m : multi) {
          const int64_t xr = std::max<int64_t>(x, c0 - (m.first - 1));
          band_steps += (int64_t)m.second * (((c1 - xr + 63) + 63) / 64 * 64);
        }
        items.push_back({band_steps, it});
        for (int32_t si : bn.segs) p->cells += (int64_t)lens[segs[(size_t)si].read] * (c1 - c0) * strands_counted;
        p->computed += BRP * band_steps * 2;
        max_w = std::max(max_w, c1 - x);
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<int64_t, WorkItem>& a, const std::pair<int64_t, WorkItem>& b) {
                       return a.first > b.first;
                     });
    LaunchGroup g;
    g.S = SP;
    g.pack = true;
    g.rows = pack_rows;
    g.item0 = (int32_t)hv.size();
    g.n_items = (int32_t)items.size();
    g.grid = (int)std::max<int64_t>(1, std::min<int64_t>(slots_pack, (int64_t)items.size()));
    if (o.waves > 0) g.grid = std::min(g.grid, o.waves);
    g.renorm_k = drift * (float)SP;
    g.scratch_wg0 = scratch_rows;
    scratch_rows += (int64_t)g.grid * pack_rows;
    max_grid = std::max(max_grid, g.grid);
    for (auto& pr : items) hv.push_back(pr.second);
    // the packed launch goes first: its items are the longest
    p->groups.insert(p->groups.begin(), g);
  }
  p->n_packs = (int32_t)bins.size();
  p->n_packed_reads = (int32_t)segs.size();
  p->n_items = (int32_t)hv.size();
  if (m3 && multiband) {
    // whole-strand band scratch rows: fewer resident waves if the rows of a
    // full grid would not fit the scratch budget
    const int64_t epc = (p->precision == DTWAX_F32) ? p->nstr : 1;
    const int64_t row_bytes = (COL_PADL + max_w + COL_PADR) * epc * 4;
    const int64_t cap = std::max<int64_t>(1, scratch_budget() / (row_bytes * (int64_t)p->groups.size()));
    scratch_rows = 0;
    max_grid = 1;
    for (auto& g : p->groups) {
      g.grid = (int)std::min<int64_t>(g.grid, cap);
      g.scratch_wg0 = scratch_rows;
      scratch_rows += g.grid;
      max_grid = std::max(max_grid, g.grid);
    }
  }
  p->grid = max_grid;
  if (dry) {
    *out = p;
    return DTWAX_OK;
  }

  auto cleanup = [&](const std::string& m) {
    dtwax_plan_destroy(p);
    return fail(DTWAX_ENOMEM, m);
  };
  if (hipMalloc(&p->d_items, std::max<size_t>(1, hv.size()) * sizeof(WorkItem)) != hipSuccess)
    return cleanup("item allocation failed");
  if (hipMalloc(&p->d_counter, p->groups.size() * sizeof(int32_t) + sizeof(int32_t)) != hipSuccess)
    return cleanup("counter allocation failed");
  if (hipMalloc(&p->d_keys, std::max<size_t>(1, (size_t)n_reads * 2) * sizeof(unsigned long long)) != hipSuccess)
    return cleanup("key allocation failed");
  if (!hv.empty() && hipMemcpy(p->d_items, hv.data(), hv.size() * sizeof(WorkItem), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup("item upload failed");
  if (!hlanes.empty()) {
    if (hipMalloc(&p->d_pack_lanes, hlanes.size() * sizeof(PackLane)) != hipSuccess ||
        hipMalloc(&p->d_pack_reads, std::max<size_t>(1, hreads.size()) * sizeof(PackRead)) != hipSuccess)
      return cleanup("pack table allocation failed");
    if (hipMemcpy(p->d_pack_lanes, hlanes.data(), hlanes.size() * sizeof(PackLane), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (!hreads.empty() && hipMemcpy(p->d_pack_reads, hreads.data(), hreads.size() * sizeof(PackRead),
                                      hipMemcpyHostToDevice) != hipSuccess))
      return cleanup("pack table upload failed");
  }
  if (p->groups.size() > 1) {
    if (hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess)
      return cleanup("event creation failed");
    for (size_t g = 1; g < p->groups.size(); g++)
      // non-blocking: no implicit ordering with a NULL caller stream, only
      // the fork/join events
      if (hipStreamCreateWithFlags(&p->groups[g].stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&p->groups[g].done, hipEventDisableTiming) != hipSuccess)
        return cleanup("stream creation failed");
  }
  if (multiband) {
    const int64_t epc = (p->precision == DTWAX_F32) ? p->nstr : 1;  // 4-byte elements per column
    p->scratch_stride = COL_PADL + max_w + COL_PADR;
    const size_t bytes = (size_t)scratch_rows * (size_t)p->scratch_stride * (size_t)epc * 4;
    if (hipMalloc(&p->d_scratch, bytes) != hipSuccess) return cleanup("band scratch allocation failed");
    if (hipMemset(p->d_scratch, 0, bytes) != hipSuccess) return cleanup("band scratch clear failed");
  }
  *out = p;
  return DTWAX_OK;
}

extern "C" {

void dtwax_plan_destroy(dtwax_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipDeviceSynchronize();
  for (void* q : {(void*)p->d_items, (void*)p->d_counter, (void*)p->d_keys, p->d_scratch, (void*)p->d_pack_lanes,
                  (void*)p->d_pack_reads})
    if (q) (void)hipFree(q);
  for (auto& g : p->groups) {
    if (g.stream) (void)hipStreamDestroy(g.stream);
    if (g.done) (void)hipEventDestroy(g.done);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  delete p;
}

int dtwax_plan_info(const dtwax_plan* p, int64_t* cells, int64_t* computed, int32_t* n_items, int32_t* segment,
                    int32_t* ref_chunk, int32_t* grid) {
  if (!p) return fail(DTWAX_EINVAL, "plan is NULL");
  if (cells) *cells = p->cells;
  if (computed) *computed = p->computed;
  if (n_items) *n_items = p->n_items;
  if (segment) *segment = p->S;
  if (ref_chunk) *ref_chunk = p->chunk;
  if (grid) *grid = p->grid;
  return DTWAX_OK;
}

int dtwax_plan_set_timing(dtwax_plan* p, int32_t enable) {
  if (!p) return fail(DTWAX_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(p->ctx->device));
  if (enable && !p->ev0) {
    HIP_TRY(hipEventCreate(&p->ev0));
    HIP_TRY(hipEventCreate(&p->ev1));
  }
  p->timing = enable != 0;
  return DTWAX_OK;
}

int dtwax_plan_packing(const dtwax_plan* p, int32_t* n_packs, int32_t* n_packed_reads) {
  if (!p) return fail(DTWAX_EINVAL, "NULL plan");
  if (n_packs) *n_packs = p->n_packs;
  if (n_packed_reads) *n_packed_reads = p->n_packed_reads;
  return DTWAX_OK;
}

int dtwax_plan_kernel_ms(dtwax_plan* p, float* ms) {
  if (!p || !ms) return fail(DTWAX_EINVAL, "NULL argument");
  if (!p->timing || !p->timed_once) return fail(DTWAX_ENOTREADY, "no timed launch yet");
  HIP_TRY(hipEventSynchronize(p->ev1));
  HIP_TRY(hipEventElapsedTime(ms, p->ev0, p->ev1));
  return DTWAX_OK;
}

}  // extern "C"

namespace dtwax {
int run_score_launch(dtwax_ctx* c, const ScoreLaunch& L, hipStream_t st) {
  HIP_TRY(hipSetDevice(c->device));
  LaunchGroup single;
  single.S = L.S;
  single.item0 = 0;
  single.n_items = L.n_items;
  single.grid = L.grid;
  single.renorm_k = L.renorm_k;
  const LaunchGroup* groups = L.n_groups ? L.groups : &single;
  const int n_groups = L.n_groups ? L.n_groups : 1;
  HIP_TRY(launch_reset(L.counter, n_groups, L.keys, (int64_t)L.n_reads * 2, st));
  KParams k{};
  if (L.ref) {
    k.ref = L.ref;
    k.n_cols = (int32_t)L.n_cols;
  } else {
    k.ref = (L.precision == DTWAX_F32)
                ? (const void*)((const float*)c->ref32_alloc + (int64_t)COL_PADL * c->n_strands)
                : (const void*)((const uint32_t*)c->ref16_alloc + COL_PADL);
    k.n_cols = (int32_t)c->n_cols;
  }
  k.q_stride = L.q_stride;
  k.last_row = L.last_row;
  k.lr_stride = L.lr_stride;
  k.lr_ntiles = L.lr_ntiles;
  k.lr_last = L.lr_last;
  k.ru_renorm = L.ru_renorm;
  k.n_items = L.n_items;
  k.queries = L.queries;
  k.q_off = L.q_off;
  k.q_len = L.q_len;
  k.items = L.items;
  k.counter = L.counter;
  k.keys = L.keys;
  k.scratch = L.scratch;
  k.scratch_stride = L.scratch_stride;
  k.state = L.state;
  k.row_stride = L.row_stride;
  if (L.ev0) HIP_TRY(hipEventRecord(L.ev0, st));
  // Groups of different segment widths run concurrently: group 0 on the
  // caller's stream, the others on their own streams forked from it and
  // joined back, so a persistent launch that drains early leaves its compute
  // units to the next width instead of idling them.
  if (n_groups > 1) HIP_TRY(hipEventRecord(L.fork, st));
  const int64_t epc = (L.precision == DTWAX_F32) ? L.nstr : 1;
  // progress-ordered waves when some SIMD runs two or more of this
  // launch's waves; one setting for every group, so that groups running
  // side by side share the SIMDs on equal terms
  bool prio = false;
  for (int g = 0; g < n_groups; g++)
    if (groups[g].n_items > 0 && std::min(groups[g].grid, groups[g].n_items) > 4 * c->n_cu) prio = true;
  k.prio_by_progress = (L.prio_ok && prio) ? 1 : 0;
  for (int g = 0; g < n_groups; g++) {
    const LaunchGroup& G = groups[g];
    hipStream_t gs = (g == 0 || !G.stream) ? st : G.stream;
    if (gs != st) HIP_TRY(hipStreamWaitEvent(gs, L.fork, 0));
    if (G.n_items > 0) {
      k.items = L.items ? L.items + G.item0 : nullptr;
      k.n_items = G.n_items;
      k.counter = L.counter + g;
      k.renorm_k = G.renorm_k;
      k.renorm_split = G.renorm_split ? 1 : 0;
      k.scratch = L.scratch ? (void*)((float*)L.scratch + G.scratch_wg0 * L.scratch_stride * epc) : nullptr;
      k.scratch_rows = G.rows;
      k.pack_lanes = L.pack_lanes;
      k.pack_reads = L.pack_reads;
      const int grid = std::max(1, std::min(G.grid, G.n_items));
      if (L.fixed) HIP_TRY(launch_sdtw_fixed(k, L.precision, grid, L.quirks, gs));
      else HIP_TRY(launch_sdtw(k, L.precision, G.S, L.nstr, grid, L.m3, gs, G.pack));
    }
    if (gs != st) {
      HIP_TRY(hipEventRecord(G.done, gs));
      HIP_TRY(hipStreamWaitEvent(st, G.done, 0));
    }
  }
  if (L.ev1) HIP_TRY(hipEventRecord(L.ev1, st));
  if (L.dists) HIP_TRY(launch_finalize_dists(L.keys, L.n_reads, L.dists_h2, L.dists, st));
  else HIP_TRY(launch_finalize(L.keys, L.n_reads, L.precision == DTWAX_F32 ? L.nstr : 2, L.out, st));
  return DTWAX_OK;
}
}  // namespace dtwax

extern "C" {

static int run_plan(dtwax_plan* p, const float* d_q, const int64_t* d_off, const int32_t* d_len,
                    dtwax_result* d_out, hipStream_t st) {
  ScoreLaunch L;
  L.precision = p->precision;
  L.S = p->S;
  L.nstr = p->nstr;
  L.grid = p->grid;
  L.queries = d_q;
  L.q_off = d_off;
  L.q_len = d_len;
  L.items = p->d_items;
  L.n_items = p->n_items;
  L.n_reads = p->n_reads;
  L.counter = p->d_counter;
  L.keys = p->d_keys;
  L.scratch = p->d_scratch;
  L.scratch_stride = p->scratch_stride;
  L.renorm_k = p->renorm_k;
  L.groups = p->groups.data();
  L.n_groups = (int)p->groups.size();
  L.pack_lanes = p->d_pack_lanes;
  L.pack_reads = p->d_pack_reads;
  L.fork = p->fork;
  L.m3 = p->m3;
  L.out = reinterpret_cast<DeviceResult*>(d_out);
  if (p->timing) {
    L.ev0 = p->ev0;
    L.ev1 = p->ev1;
    p->timed_once = p->n_items > 0;
  }
  return run_score_launch(p->ctx, L, st);
}

int dtwax_score_plan(dtwax_plan* p, const float* d_q, const int64_t* d_off, const int32_t* d_len,
                     dtwax_result* d_out, void* st) {
  if (!p || (p->n_reads && (!d_q || !d_off || !d_len || !d_out))) return fail(DTWAX_EINVAL, "NULL argument");
  if (!p->ctx->n_cols) return fail(DTWAX_ENOTREADY, "no reference uploaded");
  return run_plan(p, d_q, d_off, d_len, d_out, S_(st));
}

int dtwax_score(dtwax_ctx* c, const float* d_q, const int64_t* d_off, const int32_t* d_len, int32_t n_reads,
                const dtwax_score_opts* opts, dtwax_result* d_out, void* st) {
  if (!c) return fail(DTWAX_EINVAL, "ctx is NULL");
  if (n_reads < 0) return fail(DTWAX_EINVAL, "n_reads < 0");
  if (n_reads == 0) return DTWAX_OK;
  if (!d_q || !d_off || !d_len || !d_out) return fail(DTWAX_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(c->device));
  std::vector<int32_t> lens((size_t)n_reads);
  HIP_TRY(hipMemcpyAsync(lens.data(), d_len, (size_t)n_reads * 4, hipMemcpyDeviceToHost, S_(st)));
  HIP_TRY(hipStreamSynchronize(S_(st)));
  dtwax_plan* p = nullptr;
  int rc = dtwax_plan_create(c, lens.data(), n_reads, opts, &p);
  if (rc) return rc;
  rc = dtwax_score_plan(p, d_q, d_off, d_len, d_out, st);
  if (rc == DTWAX_OK) {
    hipError_t e = hipStreamSynchronize(S_(st));
    if (e != hipSuccess) rc = fail(DTWAX_EDEVICE, std::string("score: ") + hipGetErrorString(e));
  }
  dtwax_plan_destroy(p);
  return rc;
}

// ---- reference-layout entry points ----------------------------------------
// Geometry of the reference builds (src/include/common.hpp:60-98): QUERY_LEN
// 2048 rows; REF_LEN 48128 half2 {fwd, rev} columns (-DFP16) or 96256 float
// columns of the concatenated fwd | rev squiggle (FP32 build); tiles of 832
// columns, of which REF_LEN / 832 are scored (57 and 115); the last tile's
// last-row entries sit at REF_LEN - 832 (DTW.cu:451, 546).
namespace {
constexpr int32_t kRefQuery = 2048;
struct RefBuild {
  bool fp16;
  int32_t ref_len, n_tiles, n_cols;
};
constexpr RefBuild kBuildF16{true, 48128, 48128 / TILE_W, (48128 / TILE_W) * TILE_W};
constexpr RefBuild kBuildF32{false, 96256, 96256 / TILE_W, (96256 / TILE_W) * TILE_W};

uint32_t pad_word(const RefBuild& g) {
  const uint16_t ph = __builtin_bit_cast(uint16_t, kPadF16);
  return g.fp16 ? ((uint32_t)ph | ((uint32_t)ph << 16)) : __builtin_bit_cast(uint32_t, kPadF32);
}

int check_distances_args(dtwax_ctx* c, const void* ref, const void* query, void* dists, int32_t num_entries,
                         int32_t compat_quirks) {
  if (!c || !ref || !query || !dists) return fail(DTWAX_EINVAL, "NULL argument");
  if (num_entries < 0) return fail(DTWAX_EINVAL, "num_entries < 0");
  if (compat_quirks < 0 || compat_quirks > 2) return fail(DTWAX_EINVAL, "compat_quirks must be 0, 1 or 2");
  return DTWAX_OK;
}

// One reference-layout batch on the given buffers: rbuf = the padded
// reference buffer (pads already set), counter/keys for num_entries reads.
// Kernels only: nothing here allocates.
int distances_run(dtwax_ctx* c, const RefBuild& g, uint32_t* rbuf, int32_t* counter, unsigned long long* keys,
                  const void* ref, const void* query, void* dists, int32_t num_entries, hipStream_t st,
                  void* device_last_row, int32_t compat_quirks) {
  uint32_t* flat = rbuf + COL_PADL;
  HIP_TRY(launch_untile_reference((const uint32_t*)ref, g.n_tiles, flat, st));
  const int32_t lr_last = g.ref_len - TILE_W;
  if (compat_quirks == 2) {
    // the reference's tile-by-tile schedule: the same bits as 1 from an
    // independent kernel
    HIP_TRY(launch_compat_tiles(g.fp16, flat, query, num_entries, g.n_tiles, dists, device_last_row, g.ref_len,
                                lr_last, st));
    return DTWAX_OK;
  }
  const bool quirks = compat_quirks == 1;
  const int occ = occupancy_fixed(g.fp16 ? DTWAX_F16 : DTWAX_F32, quirks);
  ScoreLaunch L;
  L.fixed = true;
  L.precision = g.fp16 ? DTWAX_F16 : DTWAX_F32;
  L.S = 32;  // 64 lanes x 32 rows = QUERY_LEN: one band, no band scratch
  L.nstr = g.fp16 ? 2 : 1;
  L.grid = std::max(1, std::min(num_entries, std::max(1, occ) * c->n_cu));
  L.queries = (const float*)query;
  L.q_stride = kRefQuery;
  L.n_items = num_entries;
  L.n_reads = num_entries;
  L.counter = counter;
  L.keys = keys;
  L.ref = flat;
  L.n_cols = g.n_cols;
  L.quirks = quirks;
  L.last_row = device_last_row;
  L.lr_stride = g.ref_len;
  L.lr_ntiles = g.n_tiles;
  L.lr_last = lr_last;
  L.dists = dists;
  L.dists_h2 = g.fp16;
  return run_score_launch(c, L, st);
}

int distances_impl(dtwax_ctx* c, const RefBuild& g, const void* ref, const void* query, void* dists,
                   int32_t num_entries, void* stream, void* device_last_row, int32_t compat_quirks) {
  int rc = check_distances_args(c, ref, query, dists, num_entries, compat_quirks);
  if (rc || num_entries == 0) return rc;
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t st = S_(stream);
  // The untiled reference in the context's own buffer for this build (the
  // reference uploaded with dtwax_reference_upload is left alone), padded
  // like it so the sweep reads past both ends without bounds checks.
  void*& rbuf = c->compat_ref[g.fp16 ? 0 : 1];
  if (!rbuf) {
    const int64_t total = COL_PADL + g.n_cols + COL_PADR;
    if (hipMalloc(&rbuf, (size_t)total * 4) != hipSuccess) {
      rbuf = nullptr;
      return fail(DTWAX_ENOMEM, "reference-layout reference buffer");
    }
    HIP_TRY(launch_fill_u32((uint32_t*)rbuf, total, pad_word(g), st));
  }
  const bool keyed = compat_quirks != 2;
  if (keyed && c->compat_cap < num_entries) {
    if (c->compat_keys) (void)hipFree(c->compat_keys);
    if (c->compat_counter) (void)hipFree(c->compat_counter);
    c->compat_keys = nullptr;
    c->compat_counter = nullptr;
    c->compat_cap = 0;
    if (hipMalloc(&c->compat_keys, (size_t)num_entries * 2 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->compat_counter, 256) != hipSuccess)
      return fail(DTWAX_ENOMEM, "reference-layout keys");
    c->compat_cap = num_entries;
  }
  return distances_run(c, g, (uint32_t*)rbuf, c->compat_counter, c->compat_keys, ref, query, dists, num_entries,
                       st, device_last_row, compat_quirks);
}
}  // namespace

int dtwax_distances(dtwax_ctx* c, const void* ref, const void* query, void* dists, int32_t num_entries,
                    uint32_t thresh, void* stream, void* device_last_row, int32_t compat_quirks) {
  (void)thresh;  // unused by the reference kernel as well (DTW.cu:152-153)
  return distances_impl(c, kBuildF16, ref, query, dists, num_entries, stream, device_last_row, compat_quirks);
}

int dtwax_distances_f32(dtwax_ctx* c, const void* ref, const void* query, void* dists, int32_t num_entries,
                        float thresh, void* stream, void* device_last_row, int32_t compat_quirks) {
  (void)thresh;
  return distances_impl(c, kBuildF32, ref, query, dists, num_entries, stream, device_last_row, compat_quirks);
}

int dtwax_distances_noalloc(dtwax_ctx* c, const void* ref, const void* query, void* dists, int32_t num_entries,
                            uint32_t thresh, void* stream, void* device_last_row, int32_t compat_quirks) {
  (void)thresh;
  int rc = check_distances_args(c, ref, query, dists, num_entries, compat_quirks);
  if (rc || num_entries == 0) return rc;
  if (num_entries > kNoAllocMaxEntries)
    return fail(DTWAX_EINVAL, "num_entries " + std::to_string(num_entries) +
                                  " exceeds the allocation-free limit of " + std::to_string(kNoAllocMaxEntries));
  void* base = nullptr;
  if (c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[c->device];
  }
  if (!base) return fail(DTWAX_ENOTREADY, "no allocation-free scratch on this device");
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t st = S_(stream);
  uint32_t* rbuf = (uint32_t*)base;
  int32_t* counter = (int32_t*)((unsigned char*)base + kNoAllocRefBytes);
  unsigned long long* keys = (unsigned long long*)((unsigned char*)base + kNoAllocRefBytes + kNoAllocCounterBytes);
  // the pads may not hold the pad value yet (the array starts zeroed): set
  // the whole buffer, the interior is untiled over it next
  HIP_TRY(launch_fill_u32(rbuf, (int64_t)kNoAllocRefWords, pad_word(kBuildF16), st));
  return distances_run(c, kBuildF16, rbuf, counter, keys, ref, query, dists, num_entries, st, device_last_row,
                       compat_quirks);
}

// ---- self test ---------------------------------------------------------------
int dtwax_selftest_lanes(int32_t* out_host) {
  if (!out_host) return fail(DTWAX_EINVAL, "NULL argument");
  int32_t* d = nullptr;
  HIP_TRY(hipMalloc(&d, 128 * 4));
  hipError_t e = launch_selftest_lanes(d, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out_host, d, 128 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(DTWAX_EDEVICE, std::string("selftest: ") + hipGetErrorString(e));
  return DTWAX_OK;
}

}  // extern "C"
