// vbz.cpp -- decoding of FAST5 raw signal compressed with ONT's VBZ filter
// (vbz_compression, HDF5 filter id 32020).
//
// The reference reads such files only through the external HDF5 plugin
// (fast5/nanopolish_fast5_io.cpp:162-176; install_hdf5_vbz.sh:43-58 pins
// vbz_compression v1.0.1) and exits when the plugin is missing.  This is a
// restatement of the published format, used when the plugin is absent:
//
//   chunk  = [uint32 LE: decoded byte count] [payload]
//   payload = zstd frame (zstd level > 0) of a stream-VByte block, or the
//             block itself (level 0)
//   block  = keys, then data bytes: key k of value i gives its byte count,
//            values little-endian, keys packed from the low bits of each
//            key byte
//   values = zig-zag codes of sample-to-sample differences (the first
//            against 0): d = (u >> 1) ^ -(u & 1), x_i = x_{i-1} + d (int16)
//
// The filter's cd_values are (version, integer_size, delta_zigzag,
// zstd_level).  Version 0 blocks use the original stream-VByte keys (2 bits,
// 1/2/3/4 bytes).  For version 1 two layouts are described for 16-bit
// samples (2-bit keys for 0/1/2/4 bytes, 1-bit keys for 1/2 bytes); the
// layout is recognised from the block: exactly one of them must account for
// every data byte, otherwise the chunk is refused (never a guess).  Version
// 0 blocks are checked the same way against their one layout.
//
// zstd is loaded at run time (libzstd.so.1, or DTWAX_ZSTD_LIB).

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dtwax.h"

namespace dtwax {
int set_error(int code, const std::string& msg);
}

namespace {

// zstd's streaming buffers (zstd.h, stable API)
struct ZInBuf {
  const void* src;
  size_t size, pos;
};
struct ZOutBuf {
  void* dst;
  size_t size, pos;
};

struct Zstd {
  bool ok = false;
  std::string err;
  size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
  unsigned long long (*frame_size)(const void*, size_t) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  const char* (*error_name)(size_t) = nullptr;
  // streaming (optional: without it a prefix is cut from the whole frame)
  void* (*create_dctx)() = nullptr;
  size_t (*free_dctx)(void*) = nullptr;
  size_t (*dctx_reset)(void*, int) = nullptr;
  size_t (*decompress_stream)(void*, ZOutBuf*, ZInBuf*) = nullptr;
};

Zstd& zstd() {
  static Zstd z = [] {
    Zstd a;
    const char* env = std::getenv("DTWAX_ZSTD_LIB");
    const char* cands[] = {env, "libzstd.so.1", "/opt/conda/lib/libzstd.so.1", "libzstd.so"};
    void* h = nullptr;
    for (const char* c : cands)
      if (c && (h = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) {
      a.err = "zstd library not found (set DTWAX_ZSTD_LIB)";
      return a;
    }
    a.decompress = reinterpret_cast<decltype(a.decompress)>(dlsym(h, "ZSTD_decompress"));
    a.frame_size = reinterpret_cast<decltype(a.frame_size)>(dlsym(h, "ZSTD_getFrameContentSize"));
    a.is_error = reinterpret_cast<decltype(a.is_error)>(dlsym(h, "ZSTD_isError"));
    a.error_name = reinterpret_cast<decltype(a.error_name)>(dlsym(h, "ZSTD_getErrorName"));
    a.create_dctx = reinterpret_cast<decltype(a.create_dctx)>(dlsym(h, "ZSTD_createDCtx"));
    a.free_dctx = reinterpret_cast<decltype(a.free_dctx)>(dlsym(h, "ZSTD_freeDCtx"));
    a.dctx_reset = reinterpret_cast<decltype(a.dctx_reset)>(dlsym(h, "ZSTD_DCtx_reset"));
    a.decompress_stream = reinterpret_cast<decltype(a.decompress_stream)>(dlsym(h, "ZSTD_decompressStream"));
    if (!a.decompress || !a.frame_size || !a.is_error || !a.error_name) {
      a.err = "zstd library lacks required symbols";
      return a;
    }
    a.ok = true;
    return a;
  }();
  return z;
}

// Stream-VByte key layouts.
struct Layout {
  int key_bits;     // 2 or 1
  uint8_t len[4];   // bytes per key value
};
constexpr Layout kLayouts[3] = {{2, {1, 2, 3, 4}}, {2, {0, 1, 2, 4}}, {1, {1, 2, 0, 0}}};

inline int key_of(const uint8_t* keys, int64_t i, int bits) {
  if (bits == 2) return (keys[i >> 2] >> ((i & 3) * 2)) & 3;
  return (keys[i >> 3] >> (i & 7)) & 1;
}

// Total data bytes the keys announce (-1 if the layout does not fit).
int64_t data_bytes(const uint8_t* blk, int64_t size, int64_t n, const Layout& L) {
  const int64_t nkey = L.key_bits == 2 ? (n + 3) / 4 : (n + 7) / 8;
  if (nkey > size) return -1;
  int64_t total = 0;
  for (int64_t i = 0; i < n; i++) total += L.len[key_of(blk, i, L.key_bits)];
  return nkey + total == size ? total : -1;
}

}  // namespace

namespace dtwax {

// One zstd frame whose content size is announced and at most max_out bytes.
bool zstd_frame(const uint8_t* src, size_t n, size_t max_out, std::vector<uint8_t>& out, std::string& err) {
  Zstd& z = zstd();
  if (!z.ok) {
    err = z.err;
    return false;
  }
  const unsigned long long fs = z.frame_size(src, n);
  if (fs > max_out) {  // also the unknown / error sentinels (the largest values)
    err = "not a zstd frame of the expected size";
    return false;
  }
  out.resize((size_t)fs);
  const size_t got = z.decompress(out.data(), out.size(), src, n);
  if (z.is_error(got)) {
    err = std::string("zstd: ") + z.error_name(got);
    return false;
  }
  if (got != fs) {
    err = "zstd frame shorter than announced";
    return false;
  }
  return true;
}

// The first `want` bytes of one zstd frame (fewer when the frame is
// shorter), decompressed in two steps so that the caller can size the
// second from the first: out receives [0, want) after each call, the
// frame's state kept in between (one object per thread).
class ZstdPrefix {
 public:
  ~ZstdPrefix() {
    if (dctx_) zstd().free_dctx(dctx_);
  }
  bool start(const uint8_t* src, size_t n, std::string& err) {
    Zstd& z = zstd();
    if (!z.ok) return err = z.err, false;
    if (!z.create_dctx || !z.dctx_reset || !z.decompress_stream) return err = "zstd lacks streaming", false;
    if (!dctx_ && !(dctx_ = z.create_dctx())) return err = "zstd context allocation failed", false;
    z.dctx_reset(dctx_, 1);  // ZSTD_reset_session_only
    in_ = {src, n, 0};
    done_ = false;
    return true;
  }
  bool fill(std::vector<uint8_t>& out, size_t want, std::string& err) {
    Zstd& z = zstd();
    const size_t have = out.size();
    if (want <= have || done_) return true;
    out.resize(want);
    ZOutBuf ob{out.data(), want, have};
    while (ob.pos < want) {
      const size_t r = z.decompress_stream(dctx_, &ob, &in_);
      if (z.is_error(r)) return err = std::string("zstd: ") + z.error_name(r), false;
      if (r == 0) {  // frame complete
        done_ = true;
        break;
      }
      if (in_.pos == in_.size && ob.pos < want) return err = "zstd frame truncated", false;
    }
    out.resize(ob.pos);
    return true;
  }

 private:
  void* dctx_ = nullptr;
  ZInBuf in_{nullptr, 0, 0};
  bool done_ = false;
};

// 1-bit-key stream-VByte block of n zig-zag 16-bit differences (the first
// against 0), which must account for every byte of the block.
bool svb16_decode(const uint8_t* blk, size_t size, int64_t n, int16_t* dst) {
  if (data_bytes(blk, (int64_t)size, n, kLayouts[2]) < 0) return false;
  const uint8_t* p = blk + (n + 7) / 8;
  uint16_t prev = 0;
  for (int64_t i = 0; i < n; i++) {
    uint16_t u = p[0];
    if ((blk[i >> 3] >> (i & 7)) & 1) {
      u = (uint16_t)(u | (uint16_t)p[1] << 8);
      p += 2;
    } else {
      p += 1;
    }
    prev = (uint16_t)(prev + (uint16_t)((u >> 1) ^ (uint16_t)(0u - (u & 1u))));
    dst[i] = (int16_t)prev;
  }
  return true;
}

// The first m of the n values of a 1-bit-key stream-VByte block of zig-zag
// 16-bit differences, from the zstd frame `src`: only the keys and the
// leading data bytes are decompressed.  Checks the keys and bytes it reads
// (not the rest of the block).
bool svb16_zstd_prefix(const uint8_t* src, size_t n_src, int64_t n, int64_t m, int16_t* dst, std::string& err) {
  thread_local ZstdPrefix zp;
  thread_local std::vector<uint8_t> buf;
  buf.clear();
  if (!zstd().decompress_stream) {  // no streaming API: the whole frame
    const size_t max_out = (size_t)(n + 7) / 8 + 2 * (size_t)n;
    std::vector<int16_t> all((size_t)n);
    if (!zstd_frame(src, n_src, max_out, buf, err)) return false;
    if (!svb16_decode(buf.data(), buf.size(), n, all.data()))
      return err = "VBZ block does not hold " + std::to_string(n) + " samples", false;
    std::memcpy(dst, all.data(), 2 * (size_t)m);
    return true;
  }
  const size_t nkey = (size_t)(n + 7) / 8;
  if (!zp.start(src, n_src, err) || !zp.fill(buf, nkey, err)) return false;
  if (buf.size() < nkey) return err = "VBZ block shorter than its keys", false;
  size_t need = nkey;
  for (int64_t i = 0; i < m; i++) need += 1 + ((buf[(size_t)(i >> 3)] >> (i & 7)) & 1);
  if (!zp.fill(buf, need, err)) return false;
  if (buf.size() < need) return err = "VBZ block shorter than its keys announce", false;
  const uint8_t* p = buf.data() + nkey;
  uint16_t prev = 0;
  for (int64_t i = 0; i < m; i++) {
    uint16_t u = p[0];
    if ((buf[(size_t)(i >> 3)] >> (i & 7)) & 1) {
      u = (uint16_t)(u | (uint16_t)p[1] << 8);
      p += 2;
    } else {
      p += 1;
    }
    prev = (uint16_t)(prev + (uint16_t)((u >> 1) ^ (uint16_t)(0u - (u & 1u))));
    dst[i] = (int16_t)prev;
  }
  return true;
}

}  // namespace dtwax

extern "C" int dtwax_vbz_decode(const void* src, int64_t src_size, const uint32_t* cd_values, int32_t n_cd_values,
                                int16_t* dst, int64_t dst_cap, int64_t* n_out) {
  using dtwax::set_error;
  if (!src || !n_out || (dst_cap > 0 && !dst) || (n_cd_values > 0 && !cd_values))
    return set_error(DTWAX_EINVAL, "NULL argument");
  // defaults of the filter when fewer values are stored
  const uint32_t version = n_cd_values > 0 ? cd_values[0] : 0;
  const uint32_t integer_size = n_cd_values > 1 ? cd_values[1] : 2;
  const uint32_t delta_zigzag = n_cd_values > 2 ? cd_values[2] : 1;
  const uint32_t zstd_level = n_cd_values > 3 ? cd_values[3] : 1;
  if (version > 1) return set_error(DTWAX_EINVAL, "VBZ: unknown format version " + std::to_string(version));
  if (integer_size != 2) return set_error(DTWAX_EINVAL, "VBZ: only 2-byte (int16) samples are decoded");
  if (src_size < 4) return set_error(DTWAX_EINVAL, "VBZ: chunk shorter than its size header");
  const uint8_t* s = static_cast<const uint8_t*>(src);
  const uint32_t orig = (uint32_t)s[0] | (uint32_t)s[1] << 8 | (uint32_t)s[2] << 16 | (uint32_t)s[3] << 24;
  if (orig % 2) return set_error(DTWAX_EINVAL, "VBZ: odd decoded size for int16 samples");
  const int64_t n = orig / 2;
  if (n > dst_cap) return set_error(DTWAX_EINVAL, "VBZ: output buffer too small");
  const uint8_t* blk = s + 4;
  int64_t blk_size = src_size - 4;
  std::vector<uint8_t> unz;
  if (zstd_level != 0) {
    Zstd& z = zstd();
    if (!z.ok) return set_error(DTWAX_ENOTREADY, "VBZ: " + z.err);
    const unsigned long long fs = z.frame_size(blk, (size_t)blk_size);
    const unsigned long long bound = (unsigned long long)(n + 3) / 4 + 4ull * (unsigned long long)n;
    if (fs > bound)  // also the unknown / error sentinels (the largest values)
      return set_error(DTWAX_EINVAL, "VBZ: not a zstd frame of a stream-VByte block");
    unz.resize((size_t)fs);
    const size_t got = z.decompress(unz.data(), unz.size(), blk, (size_t)blk_size);
    if (z.is_error(got)) return set_error(DTWAX_EINVAL, std::string("VBZ: zstd: ") + z.error_name(got));
    if (got != fs) return set_error(DTWAX_EINVAL, "VBZ: zstd frame shorter than announced");
    blk = unz.data();
    blk_size = (int64_t)got;
  }
  const Layout* L = nullptr;
  if (version == 0) {
    if (data_bytes(blk, blk_size, n, kLayouts[0]) >= 0) L = &kLayouts[0];
  } else {
    const bool a = data_bytes(blk, blk_size, n, kLayouts[1]) >= 0;
    const bool b = data_bytes(blk, blk_size, n, kLayouts[2]) >= 0;
    if (a && b) return set_error(DTWAX_EINVAL, "VBZ: the block fits two key layouts; refusing to guess");
    L = a ? &kLayouts[1] : b ? &kLayouts[2] : nullptr;
  }
  if (!L) return set_error(DTWAX_EINVAL, "VBZ: the block does not fit the stream-VByte key layout");
  const int64_t nkey = L->key_bits == 2 ? (n + 3) / 4 : (n + 7) / 8;
  const uint8_t* p = blk + nkey;
  int32_t prev = 0;
  for (int64_t i = 0; i < n; i++) {
    const int len = L->len[key_of(blk, i, L->key_bits)];
    uint32_t u = 0;
    for (int b = 0; b < len; b++) u |= (uint32_t)p[b] << (8 * b);
    p += len;
    if (delta_zigzag) {
      const int32_t d = (int32_t)((u >> 1) ^ (0u - (u & 1u)));
      prev = (int32_t)((uint32_t)prev + (uint32_t)d);
      dst[i] = (int16_t)(uint16_t)(uint32_t)prev;
    } else {
      dst[i] = (int16_t)(uint16_t)u;
    }
  }
  *n_out = n;
  return DTWAX_OK;
}
