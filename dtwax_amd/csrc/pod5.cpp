// pod5.cpp -- POD5 input, read in place without the pod5 or Arrow libraries.
//
// POD5 is ONT's successor to multi-read FAST5 (the reference reads FAST5
// only: fast5/nanopolish_fast5_io.cpp:72-194).  One file holds
//
//   signature "\x8BPOD\r\n\x1A\n" | section marker (16 bytes)
//   | embedded Arrow IPC file | zero padding to 8 | section marker | ...
//   | "FOOTER\0\0" | footer flatbuffer | int64 LE footer length
//   | section marker | signature
//
// The footer (flatbuffer table Footer: file_identifier, software,
// pod5_version, contents: [EmbeddedFile {offset, length, format,
// content_type}]) locates the Reads (content type 0) and Signal (1) tables.
// Each is an Arrow IPC file ("ARROW1\0\0", messages, footer flatbuffer,
// int32 footer length, "ARROW1") whose record batches are read in place
// through the Arrow columnar layout:
//   Signal: signal = large_list<int16> or large_binary VBZ blobs (extension
//           "minknow.vbz"), samples = uint32 sample count of the row;
//   Reads:  read_id = fixed_size_binary(16) (extension "minknow.uuid"),
//           signal = list<uint64> of signal-table row numbers, and the
//           optional channel, well, read_number, start, calibration_offset,
//           calibration_scale and num_samples.
// Other columns (dictionary-encoded strings included) are stepped over by
// their node and buffer counts.  A VBZ blob is a zstd frame of a 1-bit-key
// stream-VByte block of zig-zag 16-bit differences, the row's `samples`
// giving the count.  A window of samples decodes only the rows it touches,
// and in the row where it ends only the block's keys and leading data bytes
// (zstd streaming).  Every offset read from the file is bounds-checked.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dtwax.h"

namespace dtwax {
int set_error(int code, const std::string& msg);
bool zstd_frame(const uint8_t* src, size_t n, size_t max_out, std::vector<uint8_t>& out, std::string& err);
bool svb16_decode(const uint8_t* blk, size_t size, int64_t n, int16_t* dst);
bool svb16_zstd_prefix(const uint8_t* src, size_t n_src, int64_t n, int64_t m, int16_t* dst, std::string& err);
}  // namespace dtwax

namespace {

template <class T>
T ld(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof v);
  return v;
}

// ---- flatbuffers (read side) ----------------------------------------------
// A table at `pos`: int32 soffset to its vtable (vtable = pos - soffset);
// vtable = uint16 vtable bytes, uint16 table bytes, uint16 field offsets
// (0 = absent).  Offset fields hold a uint32 distance forward to the target.
class Table {
 public:
  Table() = default;
  Table(const uint8_t* b, size_t n, size_t pos, bool* bad) : b_(b), n_(n), pos_(pos), bad_(bad) {
    if (!in(pos, 4)) { fail(); return; }
    const int64_t vt = (int64_t)pos - ld<int32_t>(b + pos);
    if (vt < 0 || !in((size_t)vt, 4)) { fail(); return; }
    vt_ = (size_t)vt;
    vts_ = ld<uint16_t>(b + vt_);
    ts_ = ld<uint16_t>(b + vt_ + 2);
    if (vts_ < 4 || (vts_ & 1) || !in(vt_, vts_) || ts_ < 4 || !in(pos, ts_)) { fail(); return; }
    ok_ = true;
  }
  bool ok() const { return ok_; }
  size_t at(int slot) const {  // absolute position of a field, 0 if absent
    if (!ok_) return 0;
    const size_t e = 4 + 2 * (size_t)slot;
    if (e + 2 > vts_) return 0;
    const uint16_t o = ld<uint16_t>(b_ + vt_ + e);
    if (!o) return 0;
    if (o >= ts_) {
      fail();
      return 0;
    }
    return pos_ + o;
  }
  template <class T>
  T get(int slot, T dflt) const {
    const size_t p = at(slot);
    if (!p) return dflt;
    if (!in(p, sizeof(T))) {
      fail();
      return dflt;
    }
    return ld<T>(b_ + p);
  }
  size_t ref(int slot) const {  // target of an offset field, 0 if absent
    const size_t p = at(slot);
    if (!p) return 0;
    return follow(p);
  }
  Table table(int slot) const {
    const size_t t = ref(slot);
    return t ? Table(b_, n_, t, bad_) : Table();
  }
  std::string str(int slot) const {
    const size_t t = ref(slot);
    if (!t) return std::string();
    if (!in(t, 4)) return fail(), std::string();
    const uint32_t len = ld<uint32_t>(b_ + t);
    if (!in(t + 4, (size_t)len)) return fail(), std::string();
    return std::string(reinterpret_cast<const char*>(b_ + t + 4), len);
  }
  // Vector field: element data position and count (elements of elem bytes).
  uint32_t vec(int slot, size_t elem, size_t* start) const {
    *start = 0;
    const size_t t = ref(slot);
    if (!t) return 0;
    if (!in(t, 4)) return fail(), 0;
    const uint32_t c = ld<uint32_t>(b_ + t);
    if (!in(t + 4, (size_t)c * elem)) return fail(), 0;
    *start = t + 4;
    return c;
  }
  Table elem(size_t start, uint32_t i) const {  // element i of a vector of tables
    const size_t t = follow(start + 4 * (size_t)i);
    return t ? Table(b_, n_, t, bad_) : Table();
  }
  const uint8_t* base() const { return b_; }

 private:
  size_t follow(size_t p) const {
    if (!in(p, 4)) return fail(), 0;
    const uint32_t o = ld<uint32_t>(b_ + p);
    if (!o || !in(p + o, 4)) return fail(), 0;
    return p + o;
  }
  bool in(size_t p, size_t len) const { return p <= n_ && len <= n_ - p; }
  void fail() const {
    if (bad_) *bad_ = true;
  }
  const uint8_t* b_ = nullptr;
  size_t n_ = 0, pos_ = 0, vt_ = 0;
  uint16_t vts_ = 0, ts_ = 0;
  bool* bad_ = nullptr;
  bool ok_ = false;
};

// ---- Arrow IPC -------------------------------------------------------------
// Type union ids of Schema.fbs.
enum : int {
  kNull = 1, kInt = 2, kFloat = 3, kBinary = 4, kUtf8 = 5, kBool = 6, kDecimal = 7, kDate = 8,
  kTime = 9, kTimestamp = 10, kInterval = 11, kList = 12, kStruct = 13, kUnion = 14,
  kFixedSizeBinary = 15, kFixedSizeList = 16, kMap = 17, kDuration = 18, kLargeBinary = 19,
  kLargeUtf8 = 20, kLargeList = 21
};

struct Field {
  std::string name, ext;  // ext: ARROW:extension:name
  int type = 0;
  int width = 0;          // Int bit width, FixedSizeBinary byte width, float precision
  bool is_signed = false;
  bool dict = false;      // dictionary-encoded: the batch holds indices only
  std::vector<Field> kids;
};

struct Span {
  const uint8_t* p = nullptr;
  int64_t n = 0;
};

struct Batch {
  int64_t length = 0;
  std::vector<int64_t> node_len, node_nulls;
  std::vector<Span> bufs;
};

struct ArrowFile {
  std::vector<Field> fields;
  std::vector<int> node0, buf0;  // first node / buffer of each top-level column
  int n_nodes = 0, n_bufs = 0;
  std::vector<Batch> batches;
  std::string file_identifier;   // schema metadata MINKNOW:file_identifier
};

struct Err {
  std::string msg;
  bool set(const std::string& m) {
    if (msg.empty()) msg = m;
    return false;
  }
};

bool parse_field(const Table& t, Field& f, int depth, Err& e) {
  if (!t.ok()) return e.set("malformed field");
  if (depth > 32) return e.set("fields nested too deeply");
  f.name = t.str(0);
  f.type = t.get<uint8_t>(2, 0);
  const Table ty = t.table(3);
  if (f.type == kInt) {
    f.width = ty.get<int32_t>(0, 0);
    f.is_signed = ty.get<uint8_t>(1, 0) != 0;
  } else if (f.type == kFloat) {
    f.width = ty.get<int16_t>(0, 0);  // 0 half, 1 single, 2 double
  } else if (f.type == kFixedSizeBinary) {
    f.width = ty.get<int32_t>(0, 0);
  }
  f.dict = t.ref(4) != 0;
  size_t s;
  const uint32_t nk = t.vec(5, 4, &s);
  f.kids.resize(nk);
  for (uint32_t i = 0; i < nk; i++)
    if (!parse_field(t.elem(s, i), f.kids[i], depth + 1, e)) return false;
  const uint32_t nm = t.vec(6, 4, &s);
  for (uint32_t i = 0; i < nm; i++) {
    const Table kv = t.elem(s, i);
    if (kv.str(0) == "ARROW:extension:name") f.ext = kv.str(1);
  }
  return true;
}

// Nodes and buffers a column occupies in a record batch (pre-order).
bool count_layout(const Field& f, int& nodes, int& bufs, Err& e) {
  nodes++;
  if (f.dict) {
    bufs += 2;  // validity, indices
    return true;
  }
  switch (f.type) {
    case kNull: break;
    case kInt: case kFloat: case kBool: case kDecimal: case kDate: case kTime: case kTimestamp:
    case kInterval: case kDuration: case kFixedSizeBinary: bufs += 2; break;
    case kBinary: case kUtf8: case kLargeBinary: case kLargeUtf8: bufs += 3; break;
    case kList: case kLargeList: case kMap: bufs += 2; break;
    case kStruct: case kFixedSizeList: bufs += 1; break;
    default: return e.set("column '" + f.name + "': Arrow type " + std::to_string(f.type) + " is not read");
  }
  if (f.type == kList || f.type == kLargeList || f.type == kMap || f.type == kStruct || f.type == kFixedSizeList)
    for (const Field& k : f.kids)
      if (!count_layout(k, nodes, bufs, e)) return false;
  return true;
}

// Body buffer i of a record batch, decompressed into `owned` when the batch
// uses ZSTD body compression (int64 uncompressed length, -1 = stored raw).
bool body_buffer(const uint8_t* body, int64_t body_len, int64_t off, int64_t len, bool zstd,
                 std::vector<std::unique_ptr<std::vector<uint8_t>>>& owned, Span& out, Err& e) {
  if (off < 0 || len < 0 || off > body_len || len > body_len - off) return e.set("buffer outside its message body");
  out.p = body + off;
  out.n = len;
  if (!zstd || len == 0) return true;
  if (len < 8) return e.set("compressed buffer shorter than its length prefix");
  const int64_t raw = ld<int64_t>(out.p);
  if (raw == -1) {
    out.p += 8;
    out.n -= 8;
    return true;
  }
  if (raw < 0) return e.set("bad uncompressed buffer length");
  auto v = std::make_unique<std::vector<uint8_t>>();
  std::string zerr;
  if (!dtwax::zstd_frame(out.p + 8, (size_t)(len - 8), (size_t)raw, *v, zerr) || (int64_t)v->size() != raw)
    return e.set("Arrow body buffer: " + (zerr.empty() ? std::string("size mismatch") : zerr));
  out.p = v->data();
  out.n = raw;
  owned.push_back(std::move(v));
  return true;
}

bool parse_arrow(const uint8_t* f, size_t n, ArrowFile& af, std::vector<std::unique_ptr<std::vector<uint8_t>>>& owned,
                 Err& e) {
  if (n < 18 || std::memcmp(f, "ARROW1", 6) || std::memcmp(f + n - 6, "ARROW1", 6))
    return e.set("embedded table is not an Arrow IPC file");
  const int32_t flen = ld<int32_t>(f + n - 10);
  if (flen <= 0 || (size_t)flen > n - 18) return e.set("Arrow footer length out of range");
  bool bad = false;
  const size_t fpos = n - 10 - (size_t)flen;
  const uint8_t* fb = f + fpos;
  const size_t fbn = (size_t)flen;
  if (ld<uint32_t>(fb) >= fbn) return e.set("Arrow footer root out of range");
  const Table foot(fb, fbn, ld<uint32_t>(fb), &bad);
  const Table schema = foot.table(1);
  if (!schema.ok() || bad) return e.set("Arrow footer without a schema");
  if (schema.get<int16_t>(0, 0) != 0) return e.set("big-endian Arrow data is not read");
  size_t s;
  const uint32_t nf = schema.vec(1, 4, &s);
  af.fields.resize(nf);
  for (uint32_t i = 0; i < nf; i++)
    if (!parse_field(schema.elem(s, i), af.fields[i], 0, e)) return false;
  const uint32_t nm = schema.vec(2, 4, &s);
  for (uint32_t i = 0; i < nm; i++) {
    const Table kv = schema.elem(s, i);
    if (kv.str(0) == "MINKNOW:file_identifier") af.file_identifier = kv.str(1);
  }
  for (const Field& fl : af.fields) {
    af.node0.push_back(af.n_nodes);
    af.buf0.push_back(af.n_bufs);
    if (!count_layout(fl, af.n_nodes, af.n_bufs, e)) return false;
  }
  size_t bs;
  const uint32_t nb = foot.vec(3, 24, &bs);  // Block {int64 offset; int32 meta; int64 body}
  if (bad) return e.set("malformed Arrow footer");
  for (uint32_t i = 0; i < nb; i++) {
    const uint8_t* blk = fb + bs + 24 * (size_t)i;
    const int64_t off = ld<int64_t>(blk), meta = ld<int32_t>(blk + 8), body_len = ld<int64_t>(blk + 16);
    if (off < 8 || meta < 8 || body_len < 0 || (uint64_t)off > fpos || (uint64_t)meta > fpos - off ||
        (uint64_t)body_len > fpos - off - meta)
      return e.set("Arrow record batch block out of range");
    const uint8_t* m = f + off;
    size_t mpos = 4, mlen = (size_t)ld<int32_t>(m);
    if (ld<uint32_t>(m) == 0xFFFFFFFFu) {  // continuation marker, then the length
      mpos = 8;
      mlen = (size_t)ld<int32_t>(m + 4);
    }
    if (mlen < 4 || mlen > (size_t)meta - mpos) return e.set("Arrow message length out of range");
    const uint8_t* mb = m + mpos;
    if (ld<uint32_t>(mb) >= mlen) return e.set("Arrow message root out of range");
    const Table msg(mb, mlen, ld<uint32_t>(mb), &bad);
    if (msg.get<uint8_t>(1, 0) != 3) return e.set("Arrow record batch block does not hold a record batch");
    const Table rb = msg.table(2);
    if (!rb.ok() || bad) return e.set("malformed Arrow record batch");
    bool zstd = false;
    const Table comp = rb.table(3);
    if (comp.ok()) {
      const int codec = comp.get<int8_t>(0, 0);
      if (codec != 1) return e.set("Arrow body compression other than ZSTD is not read");
      zstd = true;
    }
    Batch b;
    b.length = rb.get<int64_t>(0, 0);
    size_t ns, bfs;
    const uint32_t nn = rb.vec(1, 16, &ns), nbf = rb.vec(2, 16, &bfs);
    if (bad || b.length < 0 || (int)nn != af.n_nodes || (int)nbf != af.n_bufs)
      return e.set("Arrow record batch does not match its schema");
    for (uint32_t k = 0; k < nn; k++) {
      b.node_len.push_back(ld<int64_t>(mb + ns + 16 * (size_t)k));
      b.node_nulls.push_back(ld<int64_t>(mb + ns + 16 * (size_t)k + 8));
      if (b.node_len.back() < 0 || b.node_nulls.back() < 0) return e.set("negative Arrow node length");
    }
    const uint8_t* body = f + off + meta;
    for (uint32_t k = 0; k < nbf; k++) {
      Span sp;
      if (!body_buffer(body, body_len, ld<int64_t>(mb + bfs + 16 * (size_t)k), ld<int64_t>(mb + bfs + 16 * (size_t)k + 8),
                       zstd, owned, sp, e))
        return false;
      b.bufs.push_back(sp);
    }
    // Every row or element of a column takes at least one bit of its
    // (decompressed) buffers -- a validity bit, a value or an offset -- so no
    // batch or node length may exceed their bits.  This also keeps every
    // later size product (length x width) far below 2^63.
    uint64_t buf_bytes = 0;
    for (const Span& sp : b.bufs) buf_bytes += (uint64_t)sp.n;
    const int64_t max_len = (int64_t)std::min<uint64_t>(buf_bytes, (uint64_t)1 << 56) * 8;
    if (b.length > max_len) return e.set("Arrow record batch longer than its buffers");
    for (uint32_t k = 0; k < nn; k++)
      if (b.node_len[k] > max_len || b.node_nulls[k] > b.node_len[k])
        return e.set("Arrow node longer than its record batch buffers");
    af.batches.push_back(std::move(b));
  }
  return true;
}

int find_col(const ArrowFile& af, const char* name) {
  for (size_t i = 0; i < af.fields.size(); i++)
    if (af.fields[i].name == name) return (int)i;
  return -1;
}

inline bool valid_at(const Span& v, int64_t nulls, int64_t i) {
  if (nulls == 0 || v.n == 0) return nulls == 0;
  return (v.p[i >> 3] >> (i & 7)) & 1;
}

// Fixed-width integer / float column of one batch.
struct Scalars {
  const uint8_t* p = nullptr;
  int bytes = 0, kind = 0;  // kind: 0 unsigned, 1 signed, 2 float
  Span valid;
  int64_t nulls = 0;
  bool ok() const { return p != nullptr; }
  double at(int64_t i) const {
    const uint8_t* q = p + (size_t)i * bytes;
    if (kind == 2) return bytes == 4 ? (double)ld<float>(q) : bytes == 8 ? ld<double>(q) : NAN;
    if (kind == 1) {
      switch (bytes) {
        case 1: return (double)ld<int8_t>(q);
        case 2: return (double)ld<int16_t>(q);
        case 4: return (double)ld<int32_t>(q);
        default: return (double)ld<int64_t>(q);
      }
    }
    switch (bytes) {
      case 1: return (double)ld<uint8_t>(q);
      case 2: return (double)ld<uint16_t>(q);
      case 4: return (double)ld<uint32_t>(q);
      default: return (double)ld<uint64_t>(q);
    }
  }
  int64_t iat(int64_t i) const {  // exact for 64-bit integers
    const uint8_t* q = p + (size_t)i * bytes;
    if (bytes == 8) return ld<int64_t>(q);
    return (int64_t)at(i);
  }
  bool valid_row(int64_t i) const { return valid_at(valid, nulls, i); }
};

bool scalars(const ArrowFile& af, const Batch& b, int col, Scalars& s, Err& e) {
  const Field& f = af.fields[(size_t)col];
  if (f.dict || (f.type != kInt && f.type != kFloat)) return e.set("column '" + f.name + "' is not numeric");
  int bytes = 0, kind = 0;
  if (f.type == kInt) {
    if (f.width != 8 && f.width != 16 && f.width != 32 && f.width != 64)
      return e.set("column '" + f.name + "': integer width " + std::to_string(f.width));
    bytes = f.width / 8;
    kind = f.is_signed ? 1 : 0;
  } else {
    if (f.width != 1 && f.width != 2) return e.set("column '" + f.name + "': half floats are not read");
    bytes = f.width == 1 ? 4 : 8;
    kind = 2;
  }
  const int nd = af.node0[(size_t)col], bf = af.buf0[(size_t)col];
  if (b.node_len[(size_t)nd] != b.length) return e.set("column '" + f.name + "' length differs from its batch");
  const Span& v = b.bufs[(size_t)bf + 1];
  if (b.length > v.n / bytes) return e.set("column '" + f.name + "' shorter than its batch");
  s.p = v.p;
  s.bytes = bytes;
  s.kind = kind;
  s.valid = b.bufs[(size_t)bf];
  s.nulls = b.node_nulls[(size_t)nd];
  if (s.nulls && b.length / 8 + (b.length % 8 != 0) > s.valid.n)
    return e.set("column '" + f.name + "': validity bitmap too short");
  return true;
}

// Variable-length column: list (child values) or binary (bytes).
struct Lists {
  const uint8_t* off = nullptr;
  bool off64 = false;
  Span values;        // child values buffer / binary data
  int value_bytes = 1;
  int64_t n_values = 0;  // child length (lists) or data bytes (binary)
  int64_t begin(int64_t i) const { return off64 ? ld<int64_t>(off + 8 * (size_t)i) : ld<int32_t>(off + 4 * (size_t)i); }
};

bool lists(const ArrowFile& af, const Batch& b, int col, Lists& L, Err& e, bool want_binary, int child_bits) {
  const Field& f = af.fields[(size_t)col];
  const int nd = af.node0[(size_t)col], bf = af.buf0[(size_t)col];
  if (f.dict) return e.set("column '" + f.name + "' is dictionary-encoded");
  if (b.node_len[(size_t)nd] != b.length) return e.set("column '" + f.name + "' length differs from its batch");
  if (b.node_nulls[(size_t)nd] != 0) return e.set("column '" + f.name + "' holds nulls");
  if (want_binary) {
    if (f.type != kBinary && f.type != kLargeBinary) return e.set("column '" + f.name + "' is not binary");
    L.off64 = f.type == kLargeBinary;
    L.values = b.bufs[(size_t)bf + 2];
    L.value_bytes = 1;
    L.n_values = L.values.n;
  } else {
    if ((f.type != kList && f.type != kLargeList) || f.kids.size() != 1 || f.kids[0].type != kInt ||
        f.kids[0].dict || (child_bits && f.kids[0].width != child_bits))
      return e.set("column '" + f.name + "' is not a list of integers of the expected width");
    if (f.kids[0].width != 16 && f.kids[0].width != 32 && f.kids[0].width != 64)
      return e.set("column '" + f.name + "': list element width");
    L.off64 = f.type == kLargeList;
    if (b.node_nulls[(size_t)nd + 1] != 0) return e.set("column '" + f.name + "' holds null elements");
    L.values = b.bufs[(size_t)bf + 3];
    L.value_bytes = f.kids[0].width / 8;
    L.n_values = b.node_len[(size_t)nd + 1];
    if (L.n_values > L.values.n / L.value_bytes)
      return e.set("column '" + f.name + "': values shorter than announced");
  }
  const Span& o = b.bufs[(size_t)bf + 1];
  const int ob = L.off64 ? 8 : 4;
  if (b.length > 0 && b.length >= o.n / ob) return e.set("column '" + f.name + "': offsets too short");
  L.off = o.p;
  int64_t prev = b.length > 0 ? L.begin(0) : 0;
  if (prev < 0) return e.set("column '" + f.name + "': negative offset");
  for (int64_t i = 1; i <= b.length; i++) {
    const int64_t x = L.begin(i);
    if (x < prev) return e.set("column '" + f.name + "': offsets decrease");
    prev = x;
  }
  if (prev > L.n_values) return e.set("column '" + f.name + "': offsets beyond the values");
  return true;
}

struct SigBatch {
  int64_t rows = 0;
  bool vbz = false;
  Lists sig;
  Scalars samples;
};

}  // namespace

struct dtwax_pod5 {
  int fd = -1;
  const uint8_t* map = nullptr;
  size_t size = 0;
  std::vector<std::unique_ptr<std::vector<uint8_t>>> owned;
  std::vector<SigBatch> sig;
  std::vector<int64_t> sig_first;  // rows before batch b; back() = all rows
  std::vector<dtwax_pod5_read_info> info;
  std::vector<int64_t> row_begin;  // reads + 1
  std::vector<uint64_t> rows;

  ~dtwax_pod5() {
    if (map) munmap(const_cast<uint8_t*>(map), size);
    if (fd >= 0) close(fd);
  }

  uint32_t row_samples(uint64_t row, const SigBatch** sb, int64_t* r) const {
    const size_t b = (size_t)(std::upper_bound(sig_first.begin(), sig_first.end(), (int64_t)row) - sig_first.begin()) - 1;
    *sb = &sig[b];
    *r = (int64_t)row - sig_first[b];
    return (uint32_t)sig[b].samples.iat(*r);
  }

  // Decodes the first m values of one signal row (all `samples` of them
  // when m is -1) into dst.
  bool decode_row(uint64_t row, int16_t* dst, std::vector<uint8_t>& tmp, std::string& err, int64_t m = -1) const {
    const SigBatch* sb;
    int64_t r;
    const uint32_t n = row_samples(row, &sb, &r);
    const int64_t a = sb->sig.begin(r), z = sb->sig.begin(r + 1);
    if (m < 0 || m > (int64_t)n) m = n;
    if (!sb->vbz) {
      std::memcpy(dst, sb->sig.values.p + 2 * (size_t)a, 2 * (size_t)m);  // length checked at open
      return true;
    }
    if (m < (int64_t)n) {  // a prefix: only the keys and the leading data are decompressed
      if (!dtwax::svb16_zstd_prefix(sb->sig.values.p + a, (size_t)(z - a), n, m, dst, err)) {
        err = "signal row " + std::to_string(row) + ": " + err;
        return false;
      }
      return true;
    }
    const size_t max_out = ((size_t)n + 7) / 8 + 2 * (size_t)n;
    if (!dtwax::zstd_frame(sb->sig.values.p + a, (size_t)(z - a), max_out, tmp, err)) {
      err = "signal row " + std::to_string(row) + ": " + err;
      return false;
    }
    if (!dtwax::svb16_decode(tmp.data(), tmp.size(), n, dst)) {
      err = "signal row " + std::to_string(row) + ": VBZ block does not hold " + std::to_string(n) + " samples";
      return false;
    }
    return true;
  }
};

namespace {

const uint8_t kSig[8] = {0x8B, 'P', 'O', 'D', '\r', '\n', 0x1A, '\n'};

bool open_pod5(dtwax_pod5& P, Err& e) {
  const uint8_t* f = P.map;
  const size_t n = P.size;
  if (n < 8 + 16 + 8 + 8 + 16 + 8 || std::memcmp(f, kSig, 8) || std::memcmp(f + n - 8, kSig, 8))
    return e.set("not a POD5 file (signature)");
  if (std::memcmp(f + 8, f + n - 24, 16)) return e.set("POD5 section markers differ");
  const int64_t flen = ld<int64_t>(f + n - 32);
  if (flen <= 0 || (uint64_t)flen > n - 24 - 32 - 8) return e.set("POD5 footer length out of range");
  // "FOOTER\0\0", then the flatbuffer (zero padding may follow it)
  size_t fpos = 0;
  for (size_t pad = 0; pad < 8 && !fpos; pad++) {
    const size_t m = n - 32 - pad - (size_t)flen - 8;
    if (m >= 24 && !std::memcmp(f + m, "FOOTER\0\0", 8)) fpos = m + 8;
  }
  if (!fpos) return e.set("POD5 footer magic not found");
  const uint8_t* fb = f + fpos;
  const size_t fbn = (size_t)flen;
  if (ld<uint32_t>(fb) >= fbn) return e.set("POD5 footer root out of range");
  bool bad = false;
  const Table foot(fb, fbn, ld<uint32_t>(fb), &bad);
  const std::string file_id = foot.str(0);
  size_t s;
  const uint32_t nc = foot.vec(3, 4, &s);
  int64_t off[2] = {-1, -1}, len[2] = {0, 0};
  for (uint32_t i = 0; i < nc; i++) {
    const Table ef = foot.elem(s, i);
    const int ct = ef.get<int16_t>(3, 0);
    if (ct != 0 && ct != 1) continue;
    if (ef.get<int16_t>(2, 0) != 0) return e.set("POD5 embedded table format is not Feather V2");
    off[ct] = ef.get<int64_t>(0, -1);
    len[ct] = ef.get<int64_t>(1, 0);
    if (off[ct] < 24 || len[ct] <= 0 || (uint64_t)off[ct] > fpos || (uint64_t)len[ct] > fpos - off[ct])
      return e.set("POD5 embedded table outside the file");
  }
  if (bad) return e.set("malformed POD5 footer");
  if (off[0] < 0 || off[1] < 0) return e.set("POD5 footer lacks the Reads or Signal table");

  ArrowFile reads, signal;
  if (!parse_arrow(f + off[1], (size_t)len[1], signal, P.owned, e)) return false;
  if (!parse_arrow(f + off[0], (size_t)len[0], reads, P.owned, e)) return false;
  for (const ArrowFile* a : {&reads, &signal})
    if (!file_id.empty() && !a->file_identifier.empty() && a->file_identifier != file_id)
      return e.set("POD5 table file_identifier differs from the footer's");

  // signal table
  const int c_sig = find_col(signal, "signal"), c_smp = find_col(signal, "samples");
  if (c_sig < 0 || c_smp < 0) return e.set("POD5 Signal table lacks 'signal' or 'samples'");
  const Field& sf = signal.fields[(size_t)c_sig];
  const bool vbz = sf.type == kBinary || sf.type == kLargeBinary;
  if (vbz && sf.ext != "minknow.vbz") return e.set("POD5 binary signal column is not minknow.vbz");
  P.sig_first.push_back(0);
  for (const Batch& b : signal.batches) {
    SigBatch sb;
    sb.rows = b.length;
    sb.vbz = vbz;
    if (!lists(signal, b, c_sig, sb.sig, e, vbz, 16) || !scalars(signal, b, c_smp, sb.samples, e)) return false;
    if (sb.samples.kind == 2 || sb.samples.nulls) return e.set("POD5 'samples' column must be non-null integers");
    if (!vbz && !signal.fields[(size_t)c_sig].kids[0].is_signed) return e.set("POD5 signal values must be int16");
    for (int64_t r = 0; r < b.length; r++) {
      const int64_t c = sb.samples.iat(r);
      if (c < 0 || c > 0xFFFFFFFFll) return e.set("POD5 'samples' out of range");
      if (!vbz && sb.sig.begin(r + 1) - sb.sig.begin(r) != c)
        return e.set("POD5 signal row length differs from its 'samples'");
    }
    P.sig.push_back(sb);
    P.sig_first.push_back(P.sig_first.back() + b.length);
  }
  const int64_t total_rows = P.sig_first.back();

  // reads table
  const int c_id = find_col(reads, "read_id"), c_rows = find_col(reads, "signal");
  if (c_id < 0 || c_rows < 0) return e.set("POD5 Reads table lacks 'read_id' or 'signal'");
  const Field& idf = reads.fields[(size_t)c_id];
  if (idf.type != kFixedSizeBinary || idf.width != 16 || idf.dict) return e.set("POD5 read_id is not 16-byte binary");
  const char* opt_names[] = {"channel", "well", "read_number", "start", "calibration_offset", "calibration_scale",
                             "num_samples"};
  int opt[7];
  for (int i = 0; i < 7; i++) opt[i] = find_col(reads, opt_names[i]);
  P.row_begin.push_back(0);
  for (const Batch& b : reads.batches) {
    const int nd = reads.node0[(size_t)c_id], bf = reads.buf0[(size_t)c_id];
    if (b.node_len[(size_t)nd] != b.length || b.node_nulls[(size_t)nd] || b.length > b.bufs[(size_t)bf + 1].n / 16)
      return e.set("POD5 read_id column malformed or null");
    const uint8_t* ids = b.bufs[(size_t)bf + 1].p;
    Lists rl;
    if (!lists(reads, b, c_rows, rl, e, false, 0)) return false;
    Scalars sc[7];
    for (int i = 0; i < 7; i++)
      if (opt[i] >= 0 && !scalars(reads, b, opt[i], sc[i], e)) return false;
    for (int64_t r = 0; r < b.length; r++) {
      dtwax_pod5_read_info ri;
      std::memset(&ri, 0, sizeof ri);
      static const char hx[] = "0123456789abcdef";
      char* o = ri.read_id;
      for (int k = 0; k < 16; k++) {
        if (k == 4 || k == 6 || k == 8 || k == 10) *o++ = '-';
        *o++ = hx[ids[16 * r + k] >> 4];
        *o++ = hx[ids[16 * r + k] & 15];
      }
      auto opt_i = [&](int i, int64_t d) { return sc[i].ok() && sc[i].valid_row(r) ? sc[i].iat(r) : d; };
      auto opt_f = [&](int i) { return sc[i].ok() && sc[i].valid_row(r) ? (float)sc[i].at(r) : NAN; };
      ri.channel = (int32_t)opt_i(0, -1);
      ri.well = (int32_t)opt_i(1, -1);
      ri.read_number = (int32_t)opt_i(2, -1);
      ri.start_sample = opt_i(3, 0);
      ri.calibration_offset = opt_f(4);
      ri.calibration_scale = opt_f(5);
      const int64_t a = rl.begin(r), z = rl.begin(r + 1);
      if (z - a > INT32_MAX) return e.set("POD5 read with too many signal rows");
      ri.n_rows = (int32_t)(z - a);
      int64_t ns = 0;
      for (int64_t k = a; k < z; k++) {
        const uint8_t* q = rl.values.p + (size_t)k * rl.value_bytes;
        const uint64_t row = rl.value_bytes == 8 ? ld<uint64_t>(q) : rl.value_bytes == 4 ? ld<uint32_t>(q) : ld<uint16_t>(q);
        if (row >= (uint64_t)total_rows)
          return e.set(std::string("POD5 read ") + ri.read_id + " names signal row " + std::to_string(row) +
                       " of " + std::to_string(total_rows));
        P.rows.push_back(row);
        const SigBatch* sb;
        int64_t rr;
        ns += P.row_samples(row, &sb, &rr);
      }
      ri.n_samples = ns;
      if (opt[6] >= 0 && sc[6].valid_row(r) && sc[6].iat(r) != ns)
        return e.set(std::string("POD5 read ") + ri.read_id + ": num_samples " + std::to_string(sc[6].iat(r)) +
                     " differs from its signal rows' " + std::to_string(ns));
      P.info.push_back(ri);
      P.row_begin.push_back((int64_t)P.rows.size());
    }
  }
  return true;
}

}  // namespace

extern "C" int dtwax_pod5_open(const char* path, dtwax_pod5** out) {
  using dtwax::set_error;
  if (!path || !out) return set_error(DTWAX_EINVAL, "NULL argument");
  *out = nullptr;
  auto P = std::make_unique<dtwax_pod5>();
  P->fd = open(path, O_RDONLY | O_CLOEXEC);
  if (P->fd < 0) return set_error(DTWAX_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  struct stat st;
  if (fstat(P->fd, &st) || st.st_size <= 0) return set_error(DTWAX_EIO, std::string("empty or unreadable ") + path);
  P->size = (size_t)st.st_size;
  void* m = mmap(nullptr, P->size, PROT_READ, MAP_PRIVATE, P->fd, 0);
  if (m == MAP_FAILED) return set_error(DTWAX_EIO, std::string("cannot map ") + path + ": " + std::strerror(errno));
  P->map = static_cast<const uint8_t*>(m);
  Err e;
  if (!open_pod5(*P, e)) return set_error(DTWAX_EIO, std::string(path) + ": " + e.msg);
  *out = P.release();
  return DTWAX_OK;
}

extern "C" int dtwax_pod5_close(dtwax_pod5* f) {
  delete f;
  return DTWAX_OK;
}

extern "C" int dtwax_pod5_count(const dtwax_pod5* f, int64_t* n_reads, int64_t* n_signal_rows) {
  if (!f) return dtwax::set_error(DTWAX_EINVAL, "NULL handle");
  if (n_reads) *n_reads = (int64_t)f->info.size();
  if (n_signal_rows) *n_signal_rows = f->sig_first.back();
  return DTWAX_OK;
}

extern "C" int dtwax_pod5_reads(const dtwax_pod5* f, int64_t first, int64_t n, dtwax_pod5_read_info* out) {
  using dtwax::set_error;
  if (!f || (n > 0 && !out)) return set_error(DTWAX_EINVAL, "NULL argument");
  if (first < 0 || n < 0 || first > (int64_t)f->info.size() || n > (int64_t)f->info.size() - first)
    return set_error(DTWAX_EINVAL, "read range outside the file");
  if (n) std::memcpy(out, f->info.data() + first, (size_t)n * sizeof *out);
  return DTWAX_OK;
}

extern "C" int dtwax_pod5_signal(const dtwax_pod5* f, const int64_t* reads, int64_t n, int64_t skip, int64_t take,
                                 int16_t* dst, int64_t dst_cap, int64_t* n_out, int32_t n_threads) {
  using dtwax::set_error;
  if (!f || (n > 0 && (!reads || !n_out))) return set_error(DTWAX_EINVAL, "NULL argument");
  if (n < 0 || skip < 0) return set_error(DTWAX_EINVAL, "negative count or skip");
  if (take > INT64_MAX - skip) return set_error(DTWAX_EINVAL, "skip + take overflows");
  const int64_t nr = (int64_t)f->info.size();
  std::vector<int64_t> base((size_t)n + 1, 0);
  for (int64_t j = 0; j < n; j++) {
    if (reads[j] < 0 || reads[j] >= nr) return set_error(DTWAX_EINVAL, "read index outside the file");
    const int64_t have = std::max<int64_t>(0, f->info[(size_t)reads[j]].n_samples - skip);
    const int64_t add = take >= 0 ? take : have;
    // base[j] <= dst_cap here, so this cannot overflow; a larger total is refused
    if (add > dst_cap - base[(size_t)j]) return set_error(DTWAX_EINVAL, "destination too small");
    base[(size_t)j + 1] = base[(size_t)j] + add;
  }
  if (base[(size_t)n] > 0 && !dst) return set_error(DTWAX_EINVAL, "NULL destination");
  int threads = n_threads > 0 ? n_threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n));
  std::atomic<int64_t> next{0};
  std::mutex mu;
  std::string first_err;
  auto work = [&] {
    std::vector<uint8_t> tmp;
    std::vector<int16_t> row;
    for (int64_t j; (j = next.fetch_add(1)) < n;) {
      const int64_t ri = reads[j];
      const int64_t want = take >= 0 ? take : base[(size_t)j + 1] - base[(size_t)j];
      int16_t* out = dst + base[(size_t)j];
      int64_t pos = 0, wrote = 0;
      std::string err;
      for (int64_t k = f->row_begin[(size_t)ri]; k < f->row_begin[(size_t)ri + 1] && wrote < want; k++) {
        const SigBatch* sb;
        int64_t r;
        const int64_t c = f->row_samples(f->rows[(size_t)k], &sb, &r);
        const int64_t a = std::max<int64_t>(skip + wrote, pos), z = std::min<int64_t>(pos + c, skip + want);
        if (a < z) {
          int16_t* into;
          if (a == pos && z == pos + c) {
            into = out + (pos - skip);  // whole row inside the window
          } else {
            row.resize((size_t)c);
            into = row.data();
          }
          // a row the window ends in is decoded only up to the window's end
          if (!f->decode_row(f->rows[(size_t)k], into, tmp, err, z - pos)) break;
          if (into == row.data()) std::memcpy(out + (a - skip), row.data() + (a - pos), 2 * (size_t)(z - a));
          wrote = z - skip;
        }
        pos += c;
      }
      n_out[j] = wrote;
      if (!err.empty()) {
        std::lock_guard<std::mutex> g(mu);
        if (first_err.empty()) first_err = std::string("read ") + f->info[(size_t)ri].read_id + ": " + err;
      }
    }
  };
  if (threads == 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
  }
  if (!first_err.empty()) return set_error(DTWAX_EIO, first_err);
  return DTWAX_OK;
}
