// fast5_reader.cpp -- multi-read FAST5 (HDF5) reader.
//
// Behaviour follows the reference's vendored nanopolish reader:
//   * a file is multi-read unless it has /UniqueGlobalKey/tracking_id
//     (fast5/nanopolish_fast5_io.cpp:45-57);
//   * root groups are listed by index in name order (:72-112);
//   * /<group>/Raw/Signal is read as float with no pA conversion (:137-194,
//     conversion commented out at :181-186);
//   * channel parameters come from /<group>/channel_id attributes (:279-310);
//   * a signal compressed with the VBZ filter (id 32020) is read through the
//     HDF5 plugin when it is loaded; without it the reference exits
//     (:170-173), while this reader reads the stored chunks as they are
//     (H5Dread_chunk) and decodes them with dtwax_vbz_decode; a signal it
//     cannot decode either is reported and the next read follows.

#include "fast5_reader.h"

#include <dirent.h>
#include <dlfcn.h>
#include <hdf5.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "dtwax.h"

namespace dtwax_io {

// HDF5 is loaded at run time (dlopen) rather than linked: the library is
// only needed for FAST5 input, and linking it would put its directory (with
// an older C++ runtime) on the executable's search path ahead of the one the
// HIP runtime needs.  DTWAX_HDF5_LIB overrides the library path.
namespace {
struct H5Api {
  bool ok = false;
  std::string err;
#define H5FN(name) decltype(&::name) name = nullptr;
  H5FN(H5open) H5FN(H5Eset_auto2) H5FN(H5Fopen) H5FN(H5Fclose) H5FN(H5Lexists)
  H5FN(H5Gget_info_by_name) H5FN(H5Lget_name_by_idx) H5FN(H5Gopen2) H5FN(H5Gclose)
  H5FN(H5Aexists) H5FN(H5Aopen) H5FN(H5Aread) H5FN(H5Aclose) H5FN(H5Dopen2) H5FN(H5Dclose)
  H5FN(H5Dget_space) H5FN(H5Sget_simple_extent_dims) H5FN(H5Sclose) H5FN(H5Dread)
  H5FN(H5Dget_create_plist) H5FN(H5Pget_nfilters) H5FN(H5Pget_filter2) H5FN(H5Pclose)
  // raw chunk access (HDF5 >= 1.10.5; optional: VBZ without the plugin)
  H5FN(H5Dget_num_chunks) H5FN(H5Dget_chunk_info) H5FN(H5Dread_chunk) H5FN(H5Pget_chunk)
#undef H5FN
  hid_t* native_float = nullptr;
  hid_t* native_double = nullptr;
};

H5Api& h5() {
  static H5Api api = [] {
    H5Api a;
    const char* env = std::getenv("DTWAX_HDF5_LIB");
    const char* cands[] = {env, "/opt/conda/lib/libhdf5.so.103", "libhdf5.so.103", "libhdf5.so",
                           "libhdf5_serial.so"};
    void* h = nullptr;
    for (const char* c : cands)
      if (c && (h = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) {
      a.err = "HDF5 library not found (set DTWAX_HDF5_LIB)";
      return a;
    }
    bool all = true;
#define H5LOAD(name)                                                  \
  a.name = reinterpret_cast<decltype(a.name)>(dlsym(h, #name));       \
  all = all && a.name != nullptr;
    H5LOAD(H5open) H5LOAD(H5Eset_auto2) H5LOAD(H5Fopen) H5LOAD(H5Fclose) H5LOAD(H5Lexists)
    H5LOAD(H5Gget_info_by_name) H5LOAD(H5Lget_name_by_idx) H5LOAD(H5Gopen2) H5LOAD(H5Gclose)
    H5LOAD(H5Aexists) H5LOAD(H5Aopen) H5LOAD(H5Aread) H5LOAD(H5Aclose) H5LOAD(H5Dopen2)
    H5LOAD(H5Dclose) H5LOAD(H5Dget_space) H5LOAD(H5Sget_simple_extent_dims) H5LOAD(H5Sclose)
    H5LOAD(H5Dread) H5LOAD(H5Dget_create_plist) H5LOAD(H5Pget_nfilters) H5LOAD(H5Pget_filter2)
    H5LOAD(H5Pclose)
#undef H5LOAD
    a.H5Dget_num_chunks = reinterpret_cast<decltype(a.H5Dget_num_chunks)>(dlsym(h, "H5Dget_num_chunks"));
    a.H5Dget_chunk_info = reinterpret_cast<decltype(a.H5Dget_chunk_info)>(dlsym(h, "H5Dget_chunk_info"));
    a.H5Dread_chunk = reinterpret_cast<decltype(a.H5Dread_chunk)>(dlsym(h, "H5Dread_chunk"));
    a.H5Pget_chunk = reinterpret_cast<decltype(a.H5Pget_chunk)>(dlsym(h, "H5Pget_chunk"));
    a.native_float = reinterpret_cast<hid_t*>(dlsym(h, "H5T_NATIVE_FLOAT_g"));
    a.native_double = reinterpret_cast<hid_t*>(dlsym(h, "H5T_NATIVE_DOUBLE_g"));
    if (!all || !a.native_float || !a.native_double) {
      a.err = "HDF5 library lacks required symbols";
      return a;
    }
    a.H5open();
    a.ok = true;
    return a;
  }();
  return api;
}
}  // namespace

std::vector<std::string> list_fast5(const std::string& dir, bool* opened, bool with_pod5) {
  std::vector<std::string> out;
  DIR* d = opendir(dir.c_str());
  if (opened) *opened = d != nullptr;
  if (!d) return out;
  while (struct dirent* e = readdir(d)) {
    std::string name = e->d_name;
    // extension test as generate_load_squiggle.hpp:148: text after the last '.'
    // (a name with no '.' is compared whole)
    size_t dot = name.rfind('.');
    std::string ext = name.substr(dot == std::string::npos ? 0 : dot + 1);
    if (ext == "fast5" || (with_pod5 && ext == "pod5")) out.push_back(dir + name);
  }
  closedir(d);
  return out;
}

static double read_attr_double(hid_t obj, const char* name, bool& ok) {
  double v = 0;
  if (h5().H5Aexists(obj, name) <= 0) {
    ok = false;
    return v;
  }
  hid_t a = h5().H5Aopen(obj, name, H5P_DEFAULT);
  if (a < 0) {
    ok = false;
    return v;
  }
  if (h5().H5Aread(a, *h5().native_double, &v) < 0) ok = false;
  h5().H5Aclose(a);
  return v;
}

// The dataset's filter pipeline: whether it holds the VBZ filter, and
// whether that is its only filter (with its stored values).
struct VbzInfo {
  bool vbz = false;
  bool only = false;
  unsigned cd[8] = {0, 2, 1, 1, 0, 0, 0, 0};
  size_t n_cd = 0;
  hsize_t chunk = 0;  // chunk length (1-D)
};

static VbzInfo vbz_info(hid_t dset) {
  VbzInfo v;
  hid_t plist = h5().H5Dget_create_plist(dset);
  if (plist < 0) return v;
  int n = h5().H5Pget_nfilters(plist);
  for (int i = 0; i < n; i++) {
    unsigned int flags = 0;
    size_t nelem = 8;
    unsigned cd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    char name[64];
    H5Z_filter_t id = h5().H5Pget_filter2(plist, (unsigned)i, &flags, &nelem, cd, sizeof(name), name, nullptr);
    if (id == 32020) {
      v.vbz = true;
      v.n_cd = nelem < 8 ? nelem : 8;
      for (size_t k = 0; k < v.n_cd; k++) v.cd[k] = cd[k];
    }
  }
  v.only = v.vbz && n == 1;
  if (h5().H5Pget_chunk) {
    hsize_t dims[1] = {0};
    if (h5().H5Pget_chunk(plist, 1, dims) == 1) v.chunk = dims[0];
  }
  h5().H5Pclose(plist);
  return v;
}

// Reads a 1-D int16 VBZ dataset chunk by chunk without the filter plugin.
static bool read_vbz_native(hid_t d, const VbzInfo& v, hsize_t n, std::vector<float>& out, std::string& why) {
  H5Api& a = h5();
  if (!a.H5Dget_num_chunks || !a.H5Dget_chunk_info || !a.H5Dread_chunk || !a.H5Pget_chunk) {
    why = "the HDF5 library has no raw chunk access (1.10.5 or newer needed)";
    return false;
  }
  if (!v.only || v.chunk == 0) {
    why = "the signal's filter pipeline holds more than the VBZ filter";
    return false;
  }
  hsize_t nchunks = 0;
  const hid_t space = a.H5Dget_space(d);
  struct SpaceGuard {
    hid_t s;
    ~SpaceGuard() { if (s >= 0) h5().H5Sclose(s); }
  } guard{space};
  if (space < 0 || a.H5Dget_num_chunks(d, space, &nchunks) < 0) {
    why = "cannot count the signal's chunks";
    return false;
  }
  const uint32_t cd[4] = {v.cd[0], v.cd[1], v.cd[2], v.cd[3]};
  std::vector<uint8_t> raw;
  std::vector<int16_t> dec((size_t)v.chunk);
  out.assign((size_t)n, 0.0f);
  for (hsize_t c = 0; c < nchunks; c++) {
    hsize_t off[1] = {0}, size = 0;
    unsigned mask = 0;
    haddr_t addr = 0;
    if (a.H5Dget_chunk_info(d, space, c, off, &mask, &addr, &size) < 0) {
      why = "cannot locate a chunk of the signal";
      return false;
    }
    raw.resize((size_t)size);
    uint32_t filters = 0;
    if (a.H5Dread_chunk(d, H5P_DEFAULT, off, &filters, raw.data()) < 0) {
      why = "cannot read a chunk of the signal";
      return false;
    }
    int64_t got = 0;
    if (filters & 1u) {  // the filter was skipped for this chunk: stored as is
      got = (int64_t)std::min<size_t>(raw.size() / 2, dec.size());
      std::memcpy(dec.data(), raw.data(), (size_t)got * 2);
    } else if (dtwax_vbz_decode(raw.data(), (int64_t)raw.size(), cd, (int32_t)std::min<size_t>(v.n_cd, 4),
                                dec.data(), (int64_t)dec.size(), &got) != DTWAX_OK) {
      why = dtwax_last_error();
      return false;
    }
    for (int64_t i = 0; i < got && off[0] + (hsize_t)i < n; i++) out[(size_t)(off[0] + i)] = (float)dec[(size_t)i];
  }
  return true;
}

bool read_fast5_file(const std::string& path, std::vector<Fast5Read>& out, std::string& err) {
  if (!h5().ok) {
    err = h5().err;
    return false;
  }
  h5().H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  hid_t f = h5().H5Fopen(path.c_str(), 0x0000u /* H5F_ACC_RDONLY without its H5open() call */, H5P_DEFAULT);
  if (f < 0) {
    err = "cannot open " + path;
    return false;
  }
  const bool single = h5().H5Lexists(f, "/UniqueGlobalKey", H5P_DEFAULT) > 0 &&
                      h5().H5Lexists(f, "/UniqueGlobalKey/tracking_id", H5P_DEFAULT) > 0;
  if (single) {
    err = path + ": single-read FAST5 is not supported (multi-read layout expected)";
    h5().H5Fclose(f);
    return true;
  }
  H5G_info_t info;
  if (h5().H5Gget_info_by_name(f, "/", &info, H5P_DEFAULT) < 0) {
    err = path + ": cannot list root group";
    h5().H5Fclose(f);
    return false;
  }
  for (hsize_t i = 0; i < info.nlinks; i++) {
    ssize_t len = h5().H5Lget_name_by_idx(f, "/", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
    if (len < 0) continue;
    std::string name((size_t)len, '\0');
    h5().H5Lget_name_by_idx(f, "/", H5_INDEX_NAME, H5_ITER_INC, i, &name[0], (size_t)len + 1, H5P_DEFAULT);
    if (name.rfind("read_", 0) != 0) continue;  // not a read group
    Fast5Read r;
    r.group = name;
    // channel parameters
    std::string ch = "/" + name + "/channel_id";
    if (h5().H5Lexists(f, ("/" + name).c_str(), H5P_DEFAULT) > 0 && h5().H5Lexists(f, ch.c_str(), H5P_DEFAULT) > 0) {
      hid_t g = h5().H5Gopen2(f, ch.c_str(), H5P_DEFAULT);
      if (g >= 0) {
        bool ok = true;
        r.channel.digitisation = read_attr_double(g, "digitisation", ok);
        r.channel.offset = read_attr_double(g, "offset", ok);
        r.channel.range = read_attr_double(g, "range", ok);
        r.channel.sample_rate = read_attr_double(g, "sampling_rate", ok);
        r.channel.ok = ok;
        h5().H5Gclose(g);
      }
    }
    std::string sig = "/" + name + "/Raw/Signal";
    if (h5().H5Lexists(f, ("/" + name + "/Raw").c_str(), H5P_DEFAULT) <= 0 ||
        h5().H5Lexists(f, sig.c_str(), H5P_DEFAULT) <= 0) {
      out.push_back(std::move(r));  // no signal: zero-length read (counted as short)
      continue;
    }
    hid_t d = h5().H5Dopen2(f, sig.c_str(), H5P_DEFAULT);
    if (d < 0) {
      out.push_back(std::move(r));
      continue;
    }
    hid_t sp = h5().H5Dget_space(d);
    hsize_t n = 0;
    h5().H5Sget_simple_extent_dims(sp, &n, nullptr);
    r.samples.resize((size_t)n);
    if (n && h5().H5Dread(d, *h5().native_float, H5S_ALL, H5S_ALL, H5P_DEFAULT, r.samples.data()) < 0) {
      r.samples.clear();
      const VbzInfo v = vbz_info(d);
      std::string why;
      if (v.vbz && !read_vbz_native(d, v, n, r.samples, why)) {
        r.samples.clear();
        err = path + ": the fast5 file is compressed with VBZ, the HDF5 plugin is not loaded and "
                     "the built-in decoder failed: " + why;
      }
    }
    h5().H5Sclose(sp);
    h5().H5Dclose(d);
    out.push_back(std::move(r));
  }
  h5().H5Fclose(f);
  return true;
}

}  // namespace dtwax_io
