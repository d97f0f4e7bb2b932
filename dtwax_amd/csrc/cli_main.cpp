// cli_main.cpp -- drop-in command line driver.
//
//   dtwax_main <fast5_folder/> <kmer_model> <ref.fasta> [options]
//
// Same positional arguments, read selection and stdout log as the
// reference's offline driver (src/main.cu:53-318; log format in its
// TIMERSTOP / TIMERSTOP_CUDA macros, src/include/hpc_helpers.hpp:62-100):
//   * reference: the last FASTA entry (earlier entries are printed as
//     "name : content", load_reference.hpp:164-212), reverse complement,
//     squiggle over the first REF_LEN k-mers per strand;
//   * reads: every *.fast5 in readdir order, every read group in name order
//     (and every *.pod5 in the same listing, reads in table order, named
//     "read_<uuid>"; the reference reads FAST5 only);
//     reads shorter than ADAPTER_LEN + QUERY_LEN are skipped; samples
//     [ADAPTER_LEN, ADAPTER_LEN + QUERY_LEN) are kept as raw ADC values;
//   * mean/std normalisation (eps 1e-4), binary16 sDTW of both strands;
//   * table: j, read group, QUERY_LEN, REF_LEN, fwd score, rev score.
// Defaults reproduce the reference's compile-time geometry (QUERY_LEN 2048,
// ADAPTER_LEN 1952, REF_LEN 48128 of which 57*832 = 47,424 columns are
// scored) and its launch (main.cu:92-121, 165-204): the squiggle tiled on the
// host, queries rounded to half2 {q, q} with 32 zero pad words, scored by
// dtwax_distances (dtwax_distances_f32 with --fp32-build), whose DP is the
// reference kernel's own, seam deviations included; --clean scores the same
// launch with the clean recurrence.  Options after the positionals select the
// native mode (whole reference, any query length, medmad, other precisions),
// scored by dtwax_score.
// Exit status is non-zero on any error (the reference exits on the first
// CUDA error and on a missing VBZ plugin).
//
//   dtwax_main <kmer_model> <ref.fasta> [pipe] [options]
//
// The reference's Read Until driver (src/ru_main.cu:34-326, same program
// name, chosen here when the first argument is not a directory): reads
// batches of READS x QUERY_LEN bytes (10 x 2048) from the FIFO `pipe` in the
// working directory (generate_load_squiggle.hpp:90-130; each byte is one
// unsigned sample), normalises, scores and prints one row per query without
// a header: j, j, QUERY_LEN, REF_LEN, fwd, rev.  The reference loops
// forever; here the loop ends when the writer closes the FIFO, unless
// --keep-open (then the FIFO is reopened for the next writer).

#include <sys/stat.h>

#include <cstdint>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "dtwax.h"
#include "fast5_reader.h"

namespace {

struct Options {
  std::string dir, model, fasta;
  int query_len = 2048;      // QUERY_LEN (common.hpp:65)
  int adapter_len = 1952;    // ADAPTER_LEN (common.hpp:82)
  int64_t ref_len = 48128;   // REF_LEN per strand (common.hpp:75, FP16 build)
  int64_t scored_cols = 47424;  // (REF_LEN / (26*32)) * 26*32 (common.hpp:87-88)
  bool native = false;       // all k-mers, all columns
  int precision = DTWAX_F16;
  int norm = DTWAX_NORM_MEANSTD;
  double eps = 1e-4;
  bool dump_reads = false;   // print loaded raw samples and stop (no GPU)
  int device = 0;
  bool fp32_build = false;   // the reference's FP32 build: concatenated fwd|rev, one score
  bool clean = false;        // reference launch, clean recurrence (no seam deviations)
};

void usage() {
  std::cerr << "usage: dtwax_main <fast5_folder/> <kmer_model> <ref.fasta> [--native] [--query-len N]\n"
               "       [--adapter-len N] [--precision f16|f16_fused|f16_renorm|f32] [--norm meanstd|medmad]\n"
               "       [--fp32-build] [--clean] [--device N] [--dump-reads]\n"
               "   or: dtwax_main <kmer_model> <ref.fasta> [pipe] [--reads N] [--query-len N] [--native]\n"
               "       [--precision ...] [--norm ...] [--clean] [--device N] [--keep-open]   (Read Until FIFO loop)\n";
}

bool parse(int argc, char** argv, Options& o) {
  if (argc < 4) return false;
  o.dir = argv[1];
  o.model = argv[2];
  o.fasta = argv[3];
  for (int i = 4; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << what << "\n";
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--native") o.native = true;
    else if (a == "--query-len") o.query_len = std::atoi(next("--query-len"));
    else if (a == "--adapter-len") o.adapter_len = std::atoi(next("--adapter-len"));
    else if (a == "--precision") {
      std::string p = next("--precision");
      if (p == "f16") o.precision = DTWAX_F16;
      else if (p == "f16_fused") o.precision = DTWAX_F16_FUSED;
      else if (p == "f16_renorm") o.precision = DTWAX_F16_RENORM;
      else if (p == "f32") o.precision = DTWAX_F32;
      else return false;
    } else if (a == "--norm") {
      std::string p = next("--norm");
      if (p == "meanstd") o.norm = DTWAX_NORM_MEANSTD;
      else if (p == "medmad") o.norm = DTWAX_NORM_MEDMAD;
      else return false;
    } else if (a == "--fp32-build") {
      o.fp32_build = true;
      o.precision = DTWAX_F32;
    } else if (a == "--clean") o.clean = true;
    else if (a == "--device") o.device = std::atoi(next("--device"));
    else if (a == "--dump-reads") o.dump_reads = true;
    else if (a == "--help" || a == "-h") return false;
    else if (i == 4 && a == "pipe") continue;  // tolerated like the RU driver's argv[3]
    else {
      std::cerr << "unknown option " << a << "\n";
      return false;
    }
  }
  if (o.query_len < 1 || o.adapter_len < 0) return false;
  return true;
}

// Reads of one POD5 file with at least adapter + query samples: samples
// [adapter, adapter + query) as raw ADC values, named "read_<uuid>" as a
// FAST5 read group of the same read is.  Reads that cannot be decoded are
// reported and skipped, as unreadable FAST5 reads are.
bool load_pod5(const std::string& path, int adapter, int query, std::vector<std::string>& ids,
               std::vector<float>& raw, int64_t& invalid, int64_t& valid) {
  dtwax_pod5* f = nullptr;
  if (dtwax_pod5_open(path.c_str(), &f) != DTWAX_OK) {
    std::cerr << dtwax_last_error() << std::endl;
    return false;
  }
  int64_t n = 0;
  dtwax_pod5_count(f, &n, nullptr);
  std::vector<dtwax_pod5_read_info> info((size_t)n);
  if (n) dtwax_pod5_reads(f, 0, n, info.data());
  std::vector<int64_t> sel;
  for (int64_t i = 0; i < n; i++) {
    if (info[(size_t)i].n_samples < (int64_t)adapter + query) invalid++;
    else sel.push_back(i);
  }
  std::vector<int16_t> buf(sel.size() * (size_t)query + 1);
  std::vector<int64_t> got(sel.size() + 1);
  std::vector<char> ok(sel.size(), 1);
  if (dtwax_pod5_signal(f, sel.data(), (int64_t)sel.size(), adapter, query, buf.data(), (int64_t)buf.size(),
                        got.data(), 0) != DTWAX_OK) {
    for (size_t j = 0; j < sel.size(); j++)  // one by one: skip only the reads that fail
      if (dtwax_pod5_signal(f, &sel[j], 1, adapter, query, buf.data() + j * (size_t)query, query, &got[j], 1) !=
          DTWAX_OK) {
        std::cerr << dtwax_last_error() << std::endl;
        ok[j] = 0;
      }
  }
  for (size_t j = 0; j < sel.size(); j++) {
    if (!ok[j]) continue;
    valid++;
    ids.push_back(std::string("read_") + info[(size_t)sel[j]].read_id);
    raw.insert(raw.end(), buf.begin() + (int64_t)(j * (size_t)query), buf.begin() + (int64_t)((j + 1) * (size_t)query));
  }
  dtwax_pod5_close(f);
  return true;
}

using Clock = std::chrono::steady_clock;
struct Timer {
  Clock::time_point t0 = Clock::now();
  double ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};
void timing(const char* label, double ms) { std::cout << "TIMING: " << ms << " ms (" << label << ")" << std::endl; }

// load_reference::ref_loader (load_reference.hpp:164-212)
bool load_fasta(const std::string& path, std::string& seq) {
  std::ifstream in(path);
  if (!in.good()) {
    std::cerr << "Error opening '" << path << "'. Bailing out." << std::endl;
    return false;
  }
  std::string line, name, content;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '>') {
      if (!name.empty()) {
        std::cout << name << " : " << content << std::endl;
        name.clear();
      }
      if (!line.empty()) name = line.substr(1);
      content.clear();
    } else if (!name.empty()) {
      if (line.find(' ') != std::string::npos) {
        name.clear();
        content.clear();
      } else {
        content += line;
      }
    }
  }
  if (name.empty()) {
    std::cerr << "Invalid FASTA entry in '" << path << "'. Bailing out." << std::endl;
    return false;
  }
  seq = content;
  return true;
}

// load_reference::read_kmer_model (load_reference.hpp:131-161): "KMER mean sd"
bool load_model(const std::string& path, int& k, std::vector<float>& means) {
  std::ifstream in(path);
  if (!in.good()) {
    std::cerr << "Error opening k-mer model '" << path << "'" << std::endl;
    return false;
  }
  std::string line;
  std::map<std::string, float> m;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string kmer;
    float mean = 0, sd = 0;
    if (!(ss >> kmer >> mean >> sd)) continue;
    m.insert({kmer, mean});  // first occurrence wins, as std::map::insert
  }
  if (m.empty()) {
    std::cerr << "empty k-mer model '" << path << "'" << std::endl;
    return false;
  }
  k = (int)m.begin()->first.size();
  if (k < 1 || k > 12) return false;
  means.assign((size_t)1 << (2 * k), 0.0f);
  for (auto& kv : m) {
    if ((int)kv.first.size() != k) continue;
    int code = 0;
    bool ok = true;
    for (char ch : kv.first) {
      int b = ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : ch == 'T' ? 3 : -1;
      if (b < 0) { ok = false; break; }
      code = code * 4 + b;
    }
    if (ok) means[(size_t)code] = kv.second;
  }
  return true;
}

#define CHECK(x)                                                                  \
  do {                                                                            \
    int rc_ = (x);                                                                \
    if (rc_ != DTWAX_OK) {                                                        \
      std::cerr << "dtwax error " << rc_ << " at " #x ": " << dtwax_last_error() << std::endl; \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

// float -> binary16 bits, round to nearest even (__float2half_rn semantics).
uint16_t f32_to_f16_rn(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 rounds to inf
  if (ax >= 0x38800000u) {                                     // normal half
    const uint32_t r = ax + 0xfffu + ((ax >> 13) & 1u);
    return (uint16_t)(sign | ((r - 0x38000000u) >> 13));
  }
  float a;  // subnormal half: units of 2^-24, exact scaling then RN to an integer
  std::memcpy(&a, &ax, 4);
  return (uint16_t)(sign | (uint32_t)std::nearbyint(a * 16777216.0f));
}

// The reference's launch on the mirror entry points (src/main.cu:78-121,
// 165-204, 211-271): reference built once in the tiled layout
// h[l*832 + i*32 + j] = flat[l*832 + j*26 + i] over REF_LEN / 832 tiles
// (binary16 {fwd, rev} words, load_reference.hpp:93-97, or the concatenated
// fwd | rev floats of the FP32 build, :35-37); per batch the normalised
// queries as {q, q} half2 words (or floats) with WARP_SIZE zero pad words
// (main.cu:165-179), one distances call, the dists copied back.
struct RefMirror {
  bool fp32 = false;
  int64_t ref_len = 0;  // REF_LEN: 48128 half2 words or 96256 floats
  void *d_ref = nullptr, *d_q = nullptr, *d_dists = nullptr;
  int cap = 0;
  void init(dtwax_ctx* ctx, const std::vector<float>& levels, int64_t per_strand, bool fp32_build) {
    (void)ctx;
    fp32 = fp32_build;
    ref_len = fp32 ? 2 * per_strand : per_strand;
    std::vector<uint32_t> flat((size_t)ref_len), tiled((size_t)ref_len, 0u);
    for (int64_t i = 0; i < ref_len; i++) {
      if (fp32) {
        std::memcpy(&flat[(size_t)i], &levels[(size_t)i], 4);  // fwd then rev, as in levels
      } else {
        flat[(size_t)i] = (uint32_t)f32_to_f16_rn(levels[(size_t)i]) |
                          ((uint32_t)f32_to_f16_rn(levels[(size_t)(per_strand + i)]) << 16);
      }
    }
    const int64_t tile = 26 * 32, n_tiles = ref_len / tile;
    for (int64_t l = 0; l < n_tiles; l++)
      for (int i = 0; i < 26; i++)
        for (int j = 0; j < 32; j++) tiled[(size_t)(l * tile + i * 32 + j)] = flat[(size_t)(l * tile + j * 26 + i)];
    CHECK(dtwax_malloc(&d_ref, (size_t)ref_len * 4));
    CHECK(dtwax_memcpy_h2d(d_ref, tiled.data(), (size_t)ref_len * 4, nullptr));
  }
  // q: n normalised queries of 2048 samples -> out: n x (fp32 ? 1 : 2) scores
  void score(dtwax_ctx* ctx, const float* q, int n, bool clean, std::vector<float>& out) {
    const size_t nq = (size_t)n * 2048 + 32;
    if (n > cap) {
      if (d_q) dtwax_free(d_q);
      if (d_dists) dtwax_free(d_dists);
      CHECK(dtwax_malloc(&d_q, nq * 4));
      CHECK(dtwax_malloc(&d_dists, (size_t)n * 4));
      cap = n;
    }
    std::vector<uint32_t> words(nq, 0u);
    for (size_t i = 0; i + 32 < nq; i++) {
      if (fp32) {
        std::memcpy(&words[i], &q[i], 4);
      } else {
        const uint32_t h = f32_to_f16_rn(q[i]);
        words[i] = h | (h << 16);
      }
    }
    CHECK(dtwax_memcpy_h2d(d_q, words.data(), nq * 4, nullptr));
    const int quirks = clean ? 0 : 1;
    if (fp32) CHECK(dtwax_distances_f32(ctx, d_ref, d_q, d_dists, n, 0.0f, nullptr, nullptr, quirks));
    else CHECK(dtwax_distances(ctx, d_ref, d_q, d_dists, n, 0u, nullptr, nullptr, quirks));
    std::vector<uint32_t> d((size_t)n);
    CHECK(dtwax_memcpy_d2h(d.data(), d_dists, (size_t)n * 4, nullptr));
    out.assign((size_t)n * (fp32 ? 1 : 2), 0.0f);
    for (int j = 0; j < n; j++) {
      if (fp32) {
        std::memcpy(&out[(size_t)j], &d[(size_t)j], 4);
      } else {
        out[(size_t)(2 * j)] = f16_to_f32((uint16_t)(d[(size_t)j] & 0xffffu));
        out[(size_t)(2 * j + 1)] = f16_to_f32((uint16_t)(d[(size_t)j] >> 16));
      }
    }
  }
  void release() {
    for (void* p : {d_ref, d_q, d_dists})
      if (p) dtwax_free(p);
    d_ref = d_q = d_dists = nullptr;
  }
  static float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 31) {
      const uint32_t b = sign | 0x7f800000u | (m << 13);
      std::memcpy(&v, &b, 4);
      return v;
    }
    v = e ? std::ldexp((float)(m | 0x400u), (int)e - 25) : std::ldexp((float)m, -24);
    return sign ? -v : v;
  }
};

bool is_directory(const char* p) {
  struct stat st;
  return stat(p, &st) == 0 && S_ISDIR(st.st_mode);
}

// Read Until FIFO loop (src/ru_main.cu:172-326).
int ru_main(int argc, char** argv) {
  std::string model = argv[1], fasta = argv[2], pipe = "pipe";
  int reads = 10, query_len = 2048, device = 0, precision = DTWAX_F16, norm = DTWAX_NORM_MEANSTD;
  bool native = false, keep_open = false, clean = false;
  for (int i = 3; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << a << "\n";
        std::exit(2);
      }
      return argv[++i];
    };
    if (i == 3 && a.rfind("--", 0) != 0) pipe = a;
    else if (a == "--reads") reads = std::atoi(next());
    else if (a == "--query-len") query_len = std::atoi(next());
    else if (a == "--device") device = std::atoi(next());
    else if (a == "--native") native = true;
    else if (a == "--keep-open") keep_open = true;
    else if (a == "--clean") clean = true;
    else if (a == "--precision") {
      std::string p = next();
      if (p == "f16") precision = DTWAX_F16;
      else if (p == "f16_fused") precision = DTWAX_F16_FUSED;
      else if (p == "f16_renorm") precision = DTWAX_F16_RENORM;
      else if (p == "f32") precision = DTWAX_F32;
      else { usage(); return 2; }
    } else if (a == "--norm") {
      std::string p = next();
      if (p == "meanstd") norm = DTWAX_NORM_MEANSTD;
      else if (p == "medmad") norm = DTWAX_NORM_MEDMAD;
      else { usage(); return 2; }
    } else {
      std::cerr << "unknown option " << a << "\n";
      usage();
      return 2;
    }
  }
  if (reads < 1 || query_len < 1) {
    usage();
    return 2;
  }
  // reference: as the offline driver (REF_LEN 48128 k-mers, 57 x 832 scored)
  std::string seq;
  int k = 6;
  std::vector<float> means;
  if (!load_fasta(fasta, seq) || !load_model(model, k, means)) return 1;
  const int64_t n_kmers = (int64_t)seq.size() - k + 1;
  const int64_t ref_len = native ? n_kmers : 48128;
  if (ref_len < 1) {
    std::cerr << "reference shorter than one k-mer" << std::endl;
    return 1;
  }
  std::vector<float> levels((size_t)(2 * ref_len));
  int64_t skipped = 0;
  CHECK(dtwax_squiggle_build(seq.data(), (int64_t)seq.size(), k, means.data(), ref_len, levels.data(), &skipped));
  const int64_t n_cols = native ? ref_len : std::min<int64_t>(47424, ref_len);
  dtwax_ctx* ctx = nullptr;
  CHECK(dtwax_create(device, &ctx));
  // the reference's own launch at its geometry (ru_main.cu runs the same kernel)
  const bool mirror = !native && query_len == 2048 && precision == DTWAX_F16;
  RefMirror rm;
  if (mirror) rm.init(ctx, levels, ref_len, false);
  {
    std::vector<float> two((size_t)(2 * n_cols));
    std::copy(levels.begin(), levels.begin() + n_cols, two.begin());
    std::copy(levels.begin() + ref_len, levels.begin() + ref_len + n_cols, two.begin() + n_cols);
    CHECK(dtwax_reference_upload(ctx, two.data(), 2, n_cols));
  }
  const size_t batch = (size_t)reads * (size_t)query_len;
  std::vector<int64_t> offs((size_t)reads);
  std::vector<int32_t> lens((size_t)reads, query_len);
  for (int i = 0; i < reads; i++) offs[(size_t)i] = (int64_t)i * query_len;
  void *d_raw = nullptr, *d_norm = nullptr, *d_off = nullptr, *d_len = nullptr, *d_out = nullptr;
  CHECK(dtwax_malloc(&d_raw, batch * 4));
  CHECK(dtwax_malloc(&d_norm, batch * 4));
  CHECK(dtwax_malloc(&d_off, offs.size() * 8));
  CHECK(dtwax_malloc(&d_len, lens.size() * 4));
  CHECK(dtwax_malloc(&d_out, (size_t)reads * sizeof(dtwax_result)));
  CHECK(dtwax_memcpy_h2d(d_off, offs.data(), offs.size() * 8, nullptr));
  CHECK(dtwax_memcpy_h2d(d_len, lens.data(), lens.size() * 4, nullptr));
  dtwax_score_opts opt{};
  opt.precision = precision;
  dtwax_plan* plan = nullptr;
  CHECK(dtwax_plan_create(ctx, lens.data(), reads, &opt, &plan));
  std::vector<unsigned char> bytes(batch);
  std::vector<float> raw(batch), normed(batch), sc;
  std::vector<dtwax_result> res((size_t)reads);
  for (;;) {
    std::FILE* f = std::fopen(pipe.c_str(), "rb");  // blocks until a writer opens the FIFO
    if (!f) {
      std::cerr << "Error opening '" << pipe << "'" << std::endl;
      return 1;
    }
    for (;;) {
      size_t got = std::fread(bytes.data(), 1, batch, f);
      if (got < batch) break;  // writer closed (a partial batch is dropped)
      for (size_t i = 0; i < batch; i++) raw[i] = (float)bytes[i];
      CHECK(dtwax_memcpy_h2d(d_raw, raw.data(), batch * 4, nullptr));
      CHECK(dtwax_normalize(ctx, (const float*)d_raw, (const int64_t*)d_off, (const int32_t*)d_len, reads, norm,
                            1e-4, (float*)d_norm, nullptr, nullptr));
      if (mirror) {
        CHECK(dtwax_memcpy_d2h(normed.data(), d_norm, batch * 4, nullptr));
        rm.score(ctx, normed.data(), reads, clean, sc);
        for (int j = 0; j < reads; j++) {
          res[(size_t)j].score[0] = sc[(size_t)(2 * j)];
          res[(size_t)j].score[1] = sc[(size_t)(2 * j + 1)];
        }
      } else {
        CHECK(dtwax_score_plan(plan, (const float*)d_norm, (const int64_t*)d_off, (const int32_t*)d_len,
                               (dtwax_result*)d_out, nullptr));
        CHECK(dtwax_memcpy_d2h(res.data(), d_out, res.size() * sizeof(dtwax_result), nullptr));
      }
      for (int j = 0; j < reads; j++)
        std::cout << j << "\t" << j << "\t" << query_len << "\t" << ref_len << "\t" << res[(size_t)j].score[0]
                  << "\t" << res[(size_t)j].score[1] << "\n";
      std::cout << std::flush;
    }
    std::fclose(f);
    if (!keep_open) break;
  }
  dtwax_plan_destroy(plan);
  rm.release();
  for (void* p : {d_raw, d_norm, d_off, d_len, d_out}) dtwax_free(p);
  dtwax_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 3 && !is_directory(argv[1])) return ru_main(argc, argv);
  Options o;
  if (!parse(argc, argv, o)) {
    usage();
    return 2;
  }
  std::cout << std::flush;

  // ---- reference (load_target) --------------------------------------------
  Timer t_target;
  std::string seq;
  int k = 6;
  std::vector<float> means;
  if (!load_fasta(o.fasta, seq) || !load_model(o.model, k, means)) return 1;
  const int64_t n_kmers = (int64_t)seq.size() - k + 1;
  int64_t ref_len = o.native ? n_kmers : o.ref_len;
  if (ref_len < 1) {
    std::cerr << "reference shorter than one k-mer" << std::endl;
    return 1;
  }
  std::vector<float> levels((size_t)(2 * ref_len));
  int64_t skipped = 0;
  CHECK(dtwax_squiggle_build(seq.data(), (int64_t)seq.size(), k, means.data(), ref_len, levels.data(), &skipped));
  int64_t n_cols = o.native ? ref_len : std::min<int64_t>(o.scored_cols, ref_len);
  // the reference's own launch: its geometry, the binary16 or FP32 build
  const bool mirror = !o.native && o.query_len == 2048 && ref_len == 48128 &&
                      (o.fp32_build || o.precision == DTWAX_F16);
  RefMirror rm;
  dtwax_ctx* ctx = nullptr;
  if (!o.dump_reads) {
    CHECK(dtwax_create(o.device, &ctx));
    if (mirror) {
      rm.init(ctx, levels, ref_len, o.fp32_build);
    } else if (o.fp32_build) {
      // FP32 build: one sequence fwd[0:REF_LEN) | rev[0:REF_LEN), tiles of 832
      // over the concatenation (load_reference.hpp:35-37, common.hpp:70-73)
      const int64_t total = 2 * ref_len;
      const int64_t scored = o.native ? total : (total / 832) * 832;
      std::vector<float> cat(levels.begin(), levels.begin() + scored);
      CHECK(dtwax_reference_upload(ctx, cat.data(), 1, scored));
      n_cols = scored;
    } else {
      std::vector<float> two((size_t)(2 * n_cols));
      std::copy(levels.begin(), levels.begin() + n_cols, two.begin());
      std::copy(levels.begin() + ref_len, levels.begin() + ref_len + n_cols, two.begin() + n_cols);
      CHECK(dtwax_reference_upload(ctx, two.data(), 2, n_cols));
    }
  }
  timing("load_target", t_target.ms());

  // ---- reads (load_data) ---------------------------------------------------
  Timer t_data;
  std::cout << "cuDTW:: Loading ONT fast5 reads from " << o.dir << "\n";
  bool opened = false;
  std::vector<std::string> files = dtwax_io::list_fast5(o.dir, &opened, true);
  const int64_t need = (int64_t)o.query_len + o.adapter_len;
  std::vector<std::string> ids;
  std::vector<float> raw;
  int64_t invalid = 0, valid = 0;
  for (const auto& f : files) {
    if (f.size() > 5 && f.compare(f.size() - 5, 5, ".pod5") == 0) {
      load_pod5(f, o.adapter_len, o.query_len, ids, raw, invalid, valid);
      continue;
    }
    std::vector<dtwax_io::Fast5Read> reads;
    std::string err;
    bool ok = dtwax_io::read_fast5_file(f, reads, err);
    if (!err.empty()) std::cerr << err << std::endl;
    if (!ok) continue;
    for (auto& r : reads) {
      if ((int64_t)r.samples.size() < need) {
        invalid++;
        continue;
      }
      valid++;
      ids.push_back(r.group);
      raw.insert(raw.end(), r.samples.begin() + o.adapter_len, r.samples.begin() + need);
    }
  }
  std::cout << "cuDTW:: " << files.size() << " ONT fast5 files read\n"
            << "cuDTW:: Short/invalid reads exempted:: " << invalid << std::endl;
  std::cout << "cuDTW:: Valid reads (NUM_READS) loaded :: " << valid << std::endl;
  if (o.dump_reads) {
    for (size_t j = 0; j < ids.size(); j++) {
      std::cout << ids[j];
      for (int i = 0; i < o.query_len; i++) std::cout << " " << raw[j * o.query_len + i];
      std::cout << "\n";
    }
    return 0;
  }
  const int32_t n = (int32_t)valid;
  std::vector<int64_t> offs((size_t)n);
  std::vector<int32_t> lens((size_t)n, o.query_len);
  for (int32_t i = 0; i < n; i++) offs[(size_t)i] = (int64_t)i * o.query_len;

  Timer t_norm;
  void *d_raw = nullptr, *d_norm = nullptr, *d_off = nullptr, *d_len = nullptr, *d_out = nullptr;
  const size_t nbytes = raw.size() * sizeof(float);
  CHECK(dtwax_malloc(&d_raw, nbytes));
  CHECK(dtwax_malloc(&d_norm, nbytes));
  CHECK(dtwax_malloc(&d_off, offs.size() * 8 + 8));
  CHECK(dtwax_malloc(&d_len, lens.size() * 4 + 4));
  CHECK(dtwax_malloc(&d_out, (size_t)std::max(1, n) * sizeof(dtwax_result)));
  if (n) {
    CHECK(dtwax_memcpy_h2d(d_raw, raw.data(), nbytes, nullptr));
    CHECK(dtwax_memcpy_h2d(d_off, offs.data(), offs.size() * 8, nullptr));
    CHECK(dtwax_memcpy_h2d(d_len, lens.data(), lens.size() * 4, nullptr));
    CHECK(dtwax_normalize(ctx, (const float*)d_raw, (const int64_t*)d_off, (const int32_t*)d_len, n, o.norm,
                          o.eps, (float*)d_norm, nullptr, nullptr));
    CHECK(dtwax_device_sync());
  }
  timing("normalizer_kernel", t_norm.ms());
  std::cout << "Normalized data:\n";
  timing("load_data", t_data.ms());
  Timer t_malloc;
  timing("malloc", t_malloc.ms());

  // ---- scoring (concurrent_DTW_kernel_launch) -------------------------------
  // The reference issues batches of BLOCK_NUM = 10 reads (main.cu:211-271); the
  // batch lines are kept for log compatibility, the work is one launch.
  const int block_num = 10;
  const int batch_count = n / block_num;
  std::cout << "Batch count: " << batch_count << " num_reads:" << n << "\n";
  for (int b = 0; b <= batch_count; b++) {
    int in_batch = block_num;
    if (b == batch_count) {
      if (n % block_num == 0) {
        if (batch_count != 0) break;
        in_batch = n;
      } else {
        in_batch = n % block_num;
      }
    }
    if (in_batch == 0) continue;
    std::cout << "Processing batch_id: " << b << "\n";
    std::cout << "Issuing " << in_batch << " reads (blocks) from base addr:" << (int64_t)b * block_num * o.query_len
              << " to stream_id 0\n";
  }
  std::vector<float> normed;
  if (mirror && n) {  // the reference converts the normalised reads on the host (main.cu:165-179)
    normed.resize(raw.size());
    CHECK(dtwax_memcpy_d2h(normed.data(), d_norm, nbytes, nullptr));
  }
  Timer t_dtw;
  std::vector<float> fwd((size_t)n), rev((size_t)n);
  if (n && mirror) {
    std::vector<float> sc;
    rm.score(ctx, normed.data(), n, o.clean, sc);
    for (int32_t j = 0; j < n; j++) {
      fwd[(size_t)j] = o.fp32_build ? sc[(size_t)j] : sc[(size_t)(2 * j)];
      rev[(size_t)j] = o.fp32_build ? 0.0f : sc[(size_t)(2 * j + 1)];
    }
  } else if (n) {
    std::vector<dtwax_result> res((size_t)n);
    dtwax_score_opts opt{};
    opt.precision = o.precision;
    CHECK(dtwax_score(ctx, (const float*)d_norm, (const int64_t*)d_off, (const int32_t*)d_len, n, &opt,
                      (dtwax_result*)d_out, nullptr));
    CHECK(dtwax_memcpy_d2h(res.data(), d_out, res.size() * sizeof(dtwax_result), nullptr));
    for (int32_t j = 0; j < n; j++) {
      fwd[(size_t)j] = res[(size_t)j].score[0];
      rev[(size_t)j] = res[(size_t)j].score[1];
    }
  }
  const double dtw_ms = t_dtw.ms();
  std::cout << "TIMING: " << dtw_ms << " ms " << ((o.query_len / (dtw_ms * 1e3)) * n / 10)
            << " Mbps (concurrent_DTW_kernel_launch)" << std::endl;

  // ---- table (main.cu:279-297) ------------------------------------------------
  const int64_t ref_len_printed = o.fp32_build ? 2 * ref_len : ref_len;
  if (o.fp32_build) {
    std::cout << "Read_ID\t" << "QUERY_LEN\t" << "REF_LEN\t" << "sDTW-score\n";
    for (int32_t j = 0; j < n; j++)
      std::cout << j << "\t" << ids[(size_t)j] << "\t" << o.query_len << "\t" << ref_len_printed << "\t"
                << fwd[(size_t)j] << "\n";
  } else {
    std::cout << "Read_ID\t" << "QUERY_LEN\t" << "REF_LEN\t" << "sDTW score: fwd-strand\tsDTW score: rev-strand\n";
    for (int32_t j = 0; j < n; j++)
      std::cout << j << "\t" << ids[(size_t)j] << "\t" << o.query_len << "\t" << ref_len_printed << "\t"
                << fwd[(size_t)j] << "\t" << rev[(size_t)j] << "\n";
  }
  Timer t_free;
  rm.release();
  for (void* p : {d_raw, d_norm, d_off, d_len, d_out}) dtwax_free(p);
  dtwax_destroy(ctx);
  timing("free", t_free.ms());
  return 0;
}
