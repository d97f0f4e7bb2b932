"""Test-side VBZ encoder and FAST5 writer (test infrastructure).

The encoder follows the same published description the product's decoder
restates (a size header, zstd, stream-VByte keys and little-endian data,
zig-zag differences); it is written independently, in numpy, so that the
decoder is checked against a second restatement, and against hand-encoded
blocks in tests/test_vbz.py.  zstd itself is the system library (ctypes).
The FAST5 writer uses the HDF5 C library (ctypes) and stores chunks with
H5Dwrite_chunk, so VBZ files are written without the filter plugin.
"""
import ctypes as C

import numpy as np

_zstd = None
_h5 = None


def zstd():
    global _zstd
    if _zstd is None:
        for name in ("libzstd.so.1", "/opt/conda/lib/libzstd.so.1"):
            try:
                _zstd = C.CDLL(name)
                break
            except OSError:
                continue
        _zstd.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
        _zstd.ZSTD_compress.restype = C.c_size_t
        _zstd.ZSTD_compressBound.argtypes = [C.c_size_t]
        _zstd.ZSTD_compressBound.restype = C.c_size_t
    return _zstd


def zstd_compress(data: bytes, level=1) -> bytes:
    z = zstd()
    cap = z.ZSTD_compressBound(len(data))
    out = C.create_string_buffer(cap)
    n = z.ZSTD_compress(out, cap, data, len(data), level)
    assert n <= cap
    return out.raw[:n]


def zigzag_deltas(x):
    x = np.asarray(x, np.int64)
    d = np.diff(np.concatenate([[0], x]))
    return ((d << 1) ^ (d >> 63)).astype(np.uint64)


def svb_block(u, layout):
    """layout: "1234" (2-bit keys), "0124" (2-bit keys), "12" (1-bit keys)."""
    u = np.asarray(u, np.uint64)
    if layout == "1234":
        nb = np.where(u < 1 << 8, 1, np.where(u < 1 << 16, 2, np.where(u < 1 << 24, 3, 4)))
        key = nb - 1
        bits = 2
    elif layout == "0124":
        nb = np.where(u == 0, 0, np.where(u < 1 << 8, 1, np.where(u < 1 << 16, 2, 4)))
        key = np.select([nb == 0, nb == 1, nb == 2], [0, 1, 2], 3)
        bits = 2
    else:
        assert np.all(u < 1 << 16)
        nb = np.where(u < 1 << 8, 1, 2)
        key = nb - 1
        bits = 1
    per = 8 // bits
    keys = np.zeros((u.size + per - 1) // per, np.uint8)
    for i, k in enumerate(key):
        keys[i // per] |= np.uint8(int(k) << ((i % per) * bits))
    data = bytearray()
    for v, b in zip(u, nb):
        data += int(v).to_bytes(8, "little")[: int(b)]
    return keys.tobytes() + bytes(data)


def vbz_chunk(samples, version=0, layout=None, zstd_level=1):
    x = np.asarray(samples, np.int16)
    block = svb_block(zigzag_deltas(x), layout or ("1234" if version == 0 else "0124"))
    payload = zstd_compress(block, zstd_level) if zstd_level else block
    return int(2 * x.size).to_bytes(4, "little") + payload


# ---- FAST5 writing (HDF5 C API) --------------------------------------------
def h5():
    global _h5
    if _h5 is None:
        L = C.CDLL("/opt/conda/lib/libhdf5.so")
        hid, herr = C.c_int64, C.c_int
        sig = {
            "H5open": ([], herr), "H5Fcreate": ([C.c_char_p, C.c_uint, hid, hid], hid),
            "H5Fclose": ([hid], herr),
            "H5Gcreate2": ([hid, C.c_char_p, hid, hid, hid], hid), "H5Gclose": ([hid], herr),
            "H5Screate": ([C.c_int], hid),
            "H5Screate_simple": ([C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], hid),
            "H5Sclose": ([hid], herr),
            "H5Acreate2": ([hid, C.c_char_p, hid, hid, hid, hid], hid),
            "H5Awrite": ([hid, hid, C.c_void_p], herr), "H5Aclose": ([hid], herr),
            "H5Pcreate": ([hid], hid), "H5Pclose": ([hid], herr),
            "H5Pset_chunk": ([hid, C.c_int, C.POINTER(C.c_uint64)], herr),
            "H5Pset_filter": ([hid, C.c_int, C.c_uint, C.c_size_t, C.POINTER(C.c_uint)], herr),
            "H5Dcreate2": ([hid, C.c_char_p, hid, hid, hid, hid, hid], hid),
            "H5Dwrite_chunk": ([hid, hid, C.c_uint32, C.POINTER(C.c_uint64), C.c_size_t,
                                C.c_void_p], herr),
            "H5Dwrite": ([hid, hid, hid, hid, hid, C.c_void_p], herr), "H5Dclose": ([hid], herr),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        assert L.H5open() >= 0
        _h5 = L
    return _h5


def _g(L, name):
    return C.c_int64.in_dll(L, name).value


def write_fast5(path, reads, vbz=None, chunk=4000):
    """reads: [(group name "read_<id>", int16 signal, {attr: float})].
    vbz: None (plain chunks) or dict(version=, layout=, zstd_level=)."""
    L = h5()
    f = L.H5Fcreate(str(path).encode(), 0x0002, 0, 0)  # H5F_ACC_TRUNC
    assert f >= 0
    i16 = _g(L, "H5T_STD_I16LE_g")
    f64 = _g(L, "H5T_IEEE_F64LE_g")
    native_i16 = _g(L, "H5T_NATIVE_SHORT_g")
    native_f64 = _g(L, "H5T_NATIVE_DOUBLE_g")
    dcpl_cls = _g(L, "H5P_CLS_DATASET_CREATE_ID_g")
    for name, sig, attrs in reads:
        sig = np.ascontiguousarray(sig, np.int16)
        g = L.H5Gcreate2(f, f"/{name}".encode(), 0, 0, 0)
        ch = L.H5Gcreate2(f, f"/{name}/channel_id".encode(), 0, 0, 0)
        sc = L.H5Screate(0)  # H5S_SCALAR
        for k, v in attrs.items():
            a = L.H5Acreate2(ch, k.encode(), f64, sc, 0, 0)
            val = C.c_double(v)
            assert L.H5Awrite(a, native_f64, C.byref(val)) >= 0
            L.H5Aclose(a)
        L.H5Sclose(sc)
        raw = L.H5Gcreate2(f, f"/{name}/Raw".encode(), 0, 0, 0)
        n = sig.size
        dims = (C.c_uint64 * 1)(n)
        sp = L.H5Screate_simple(1, dims, None)
        dcpl = L.H5Pcreate(dcpl_cls)
        cdim = (C.c_uint64 * 1)(min(chunk, n))
        assert L.H5Pset_chunk(dcpl, 1, cdim) >= 0
        if vbz is not None:
            cd = (C.c_uint * 4)(vbz.get("version", 0), 2, 1, vbz.get("zstd_level", 1))
            assert L.H5Pset_filter(dcpl, 32020, 1, 4, cd) >= 0  # H5Z_FLAG_OPTIONAL
        d = L.H5Dcreate2(raw, b"Signal", i16, sp, 0, dcpl, 0)
        assert d >= 0
        if vbz is None:
            assert L.H5Dwrite(d, native_i16, 0, 0, 0, sig.ctypes.data) >= 0
        else:
            c = min(chunk, n)
            for off in range(0, n, c):
                part = np.zeros(c, np.int16)  # a chunk is stored whole (fill beyond the end)
                part[: min(c, n - off)] = sig[off:off + c]
                blob = vbz_chunk(part, vbz.get("version", 0), vbz.get("layout"),
                                 vbz.get("zstd_level", 1))
                o = (C.c_uint64 * 1)(off)
                assert L.H5Dwrite_chunk(d, 0, 0, o, len(blob), blob) >= 0
        for h in (d,):
            L.H5Dclose(h)
        L.H5Pclose(dcpl)
        L.H5Sclose(sp)
        for h in (raw, ch, g):
            L.H5Gclose(h)
    L.H5Fclose(f)
