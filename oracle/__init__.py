"""CPU oracle for the sDTW classifier path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker (or, for the baseline, as
the timed CPU port).  The product (``dtwax_amd``) never imports it.

Thin ctypes wrappers over ``liboracle.so`` (built from the C files in this
directory by ``make -C oracle``).  Every C function cites the reference
file:line it restates; see oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f32_to_f16_array.argtypes = [_f32p, _u16p, C.c_int64]
        L.orc_f16_to_f32_array.argtypes = [_u16p, _f32p, C.c_int64]
        for n in ("orc_h_add", "orc_h_sub", "orc_h_mul"):
            getattr(L, n).restype = C.c_uint16
            getattr(L, n).argtypes = [C.c_uint16, C.c_uint16]
        L.orc_h_fma.restype = C.c_uint16
        L.orc_h_fma.argtypes = [C.c_uint16, C.c_uint16, C.c_uint16]
        L.orc_sdtw_f64.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.orc_sdtw_f32.argtypes = [_f32p, C.c_int64, _f32p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        for n in ("orc_sdtw_f16", "orc_sdtw_f16_fused", "orc_sdtw_f16_scalar"):
            getattr(L, n).argtypes = [_u16p, C.c_int64, _u16p, C.c_int64, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_uint16), C.POINTER(C.c_int64)]
        L.orc_sdtw_f64_trace.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, _i64p, C.c_int, _f64p]
        L.orc_sdtw_bruteforce_f64.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, _f64p]
        L.orc_batch_f64.argtypes = [_f64p, _i64p, _i64p, C.c_int64, _f64p, C.c_int64,
                                    C.c_int, _f64p, _i64p, C.c_int]
        L.orc_batch_f32.argtypes = [_f32p, _i64p, _i64p, C.c_int64, _f32p, C.c_int64,
                                    C.c_int, _f32p, _i64p, C.c_int]
        L.orc_batch_f16.argtypes = [_u16p, _i64p, _i64p, C.c_int64, _u16p, C.c_int64,
                                    C.c_int, C.c_int, _u16p, _i64p, C.c_int]
        L.orc_sdtw3_f64.argtypes = [_f64p, C.c_int64, _f64p, C.c_int64, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]
        L.orc_sdtw3_f32.argtypes = [_f32p, C.c_int64, _f32p, C.c_int64, C.POINTER(C.c_float),
                                    C.POINTER(C.c_int64)]
        L.orc_sdtw3_f16.argtypes = [_u16p, C.c_int64, _u16p, C.c_int64, C.POINTER(C.c_uint16),
                                    C.POINTER(C.c_int64)]
        L.orc_batch3_f32.argtypes = [_f32p, _i64p, _i64p, C.c_int64, _f32p, C.c_int64,
                                     C.c_int, _f32p, _i64p, C.c_int]
        L.orc_batch3_f16.argtypes = [_u16p, _i64p, _i64p, C.c_int64, _u16p, C.c_int64,
                                     C.c_int, _u16p, _i64p, C.c_int]
        L.orc_cpu_baseline_f32.restype = C.c_int
        L.orc_cpu_baseline_f32.argtypes = [_f32p, _i64p, _i64p, C.c_int64, _f32p,
                                           C.c_int64, C.c_int, _f32p, _i64p, C.c_int]
        L.orc_sdtw_f16_renorm.argtypes = [_u16p, C.c_int64, _u16p, C.c_int64, C.c_int, C.c_float,
                                          C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        L.orc_sdtw_f16_renorm_split.argtypes = [_u16p, C.c_int64, _u16p, C.c_int64, C.c_float,
                                                C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        L.orc_sdtw_f16_renorm_row.argtypes = [_u16p, C.c_int64, _u16p, C.c_int64, C.c_int, C.c_int, C.c_float,
                                              _f32p, C.POINTER(C.c_float)]
        L.orc_renorm_split_rows.restype = C.c_int64
        L.orc_renorm_split_rows.argtypes = [C.c_int64]
        L.orc_batch_f16_renorm.argtypes = [_u16p, _i64p, _i64p, C.c_int64, _u16p, C.c_int64, C.c_int,
                                           C.c_int, C.c_float, _f32p, _i64p, C.c_int]
        L.orc_emulate_reference.restype = C.c_int
        L.orc_emulate_reference.argtypes = [C.c_void_p, _f32p, _f32p, C.c_int64, _f32p, C.c_void_p,
                                            C.c_int]
        L.orc_numpy_sum.restype = C.c_double
        L.orc_numpy_sum.argtypes = [_f64p, C.c_int64]
        L.orc_python_znorm.argtypes = [_f64p, C.c_int64, _f64p]
        L.orc_norm_meanstd.argtypes = [_f32p, C.c_int64, C.c_double, _f32p,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_norm_medmad.argtypes = [_f32p, C.c_int64, _f32p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_squiggle_reference.restype = C.c_int64
        L.orc_squiggle_reference.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_int,
                                             _f32p, C.c_int64, _f32p]
        _LIB = L
    return _LIB


class EmuGeom(C.Structure):
    _fields_ = [("query_len", C.c_int), ("prefix_len", C.c_int), ("segment", C.c_int),
                ("warp", C.c_int), ("ref_len", C.c_int), ("fp16", C.c_int),
                ("fix_q1", C.c_int), ("fix_q2", C.c_int)]


# ------------------------------------------------------------------ binary16
def f32_to_f16(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.empty(x.size, np.uint16)
    lib().orc_f32_to_f16_array(x, out, x.size)
    return out


def f16_to_f32(h) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16).ravel()
    out = np.empty(h.size, np.float32)
    lib().orc_f16_to_f32_array(h, out, h.size)
    return out


# ------------------------------------------------------------- single reads
def _ptr(a):
    return None if a is None else a.ctypes.data


def sdtw_f64(q, s, init_row=None, want_row=False):
    q = np.ascontiguousarray(q, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    init = None if init_row is None else np.ascontiguousarray(init_row, np.float64)
    row = np.empty(s.size, np.float64) if want_row else None
    sc, am = C.c_double(), C.c_int64()
    lib().orc_sdtw_f64(q, q.size, s, s.size, _ptr(init), _ptr(row), C.byref(sc), C.byref(am))
    return (sc.value, am.value, row) if want_row else (sc.value, am.value)


def sdtw_f32(q, s, init_row=None, want_row=False):
    q = np.ascontiguousarray(q, np.float32)
    s = np.ascontiguousarray(s, np.float32)
    init = None if init_row is None else np.ascontiguousarray(init_row, np.float32)
    row = np.empty(s.size, np.float32) if want_row else None
    sc, am = C.c_float(), C.c_int64()
    lib().orc_sdtw_f32(q, q.size, s, s.size, _ptr(init), _ptr(row), C.byref(sc), C.byref(am))
    return (sc.value, am.value, row) if want_row else (sc.value, am.value)


def sdtw_f16(q_h, s_h, init_row=None, want_row=False, fused=False, scalar=False):
    """q_h, s_h: binary16 bit patterns (uint16).  Returns (score_bits, argmin[, row])."""
    q = np.ascontiguousarray(q_h, np.uint16)
    s = np.ascontiguousarray(s_h, np.uint16)
    init = None if init_row is None else np.ascontiguousarray(init_row, np.uint16)
    row = np.empty(s.size, np.uint16) if want_row else None
    sc, am = C.c_uint16(), C.c_int64()
    fn = (lib().orc_sdtw_f16_fused if fused else
          lib().orc_sdtw_f16_scalar if scalar else lib().orc_sdtw_f16)
    fn(q, q.size, s, s.size, _ptr(init), _ptr(row), C.byref(sc), C.byref(am))
    return (sc.value, am.value, row) if want_row else (sc.value, am.value)


def sdtw3_f64(q, s):
    """Three-move recurrence (left move too), FP64: (score, argmin)."""
    q = np.ascontiguousarray(q, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    sc, am = C.c_double(), C.c_int64()
    lib().orc_sdtw3_f64(q, q.size, s, s.size, C.byref(sc), C.byref(am))
    return sc.value, am.value


def sdtw3_f32(q, s):
    q = np.ascontiguousarray(q, np.float32)
    s = np.ascontiguousarray(s, np.float32)
    sc, am = C.c_float(), C.c_int64()
    lib().orc_sdtw3_f32(q, q.size, s, s.size, C.byref(sc), C.byref(am))
    return sc.value, am.value


def sdtw3_f16(q_h, s_h):
    """binary16 bit patterns in, (score bits, argmin) out."""
    q = np.ascontiguousarray(q_h, np.uint16)
    s = np.ascontiguousarray(s_h, np.uint16)
    sc, am = C.c_uint16(), C.c_int64()
    lib().orc_sdtw3_f16(q, q.size, s, s.size, C.byref(sc), C.byref(am))
    return sc.value, am.value


def batch3_f32(queries, strands, n_threads=0):
    strands = np.ascontiguousarray(np.atleast_2d(strands), np.float32)
    flat, offs, lens = _pack(queries, np.float32)
    n, k = len(queries), strands.shape[0]
    sc = np.empty(n * k, np.float32)
    am = np.empty(n * k, np.int64)
    lib().orc_batch3_f32(flat, offs, lens, n, strands.ravel(), strands.shape[1], k, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


def batch3_f16(queries_h, strands_h, n_threads=0):
    strands = np.ascontiguousarray(np.atleast_2d(strands_h), np.uint16)
    flat, offs, lens = _pack(queries_h, np.uint16)
    n, k = len(queries_h), strands.shape[0]
    sc = np.empty(n * k, np.uint16)
    am = np.empty(n * k, np.int64)
    lib().orc_batch3_f16(flat, offs, lens, n, strands.ravel(), strands.shape[1], k, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


def trace_f64(q, s, cols):
    q = np.ascontiguousarray(q, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    cols = np.ascontiguousarray(cols, np.int64)
    out = np.empty(q.size * cols.size, np.float64)
    lib().orc_sdtw_f64_trace(q, q.size, s, s.size, cols, cols.size, out)
    return out.reshape(q.size, cols.size)


def bruteforce_f64(q, s):
    q = np.ascontiguousarray(q, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    m = np.empty(q.size * s.size, np.float64)
    lib().orc_sdtw_bruteforce_f64(q, q.size, s, s.size, m)
    return m.reshape(q.size, s.size)


# -------------------------------------------------------------------- batches
def _pack(queries, dtype):
    lens = np.array([len(x) for x in queries], np.int64)
    offs = np.zeros(len(queries), np.int64)
    if len(queries) > 1:
        offs[1:] = np.cumsum(lens)[:-1]
    flat = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype) for x in queries])
                                if len(queries) else np.zeros(0, dtype))
    return flat, offs, lens


def batch_f64(queries, strands, n_threads=0):
    """queries: list of float arrays; strands: (n_strands, N) array."""
    strands = np.ascontiguousarray(np.atleast_2d(strands), np.float64)
    flat, offs, lens = _pack(queries, np.float64)
    n, k = len(queries), strands.shape[0]
    sc = np.empty(n * k, np.float64)
    am = np.empty(n * k, np.int64)
    lib().orc_batch_f64(flat, offs, lens, n, strands.ravel(), strands.shape[1], k, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


def batch_f32(queries, strands, n_threads=0):
    strands = np.ascontiguousarray(np.atleast_2d(strands), np.float32)
    flat, offs, lens = _pack(queries, np.float32)
    n, k = len(queries), strands.shape[0]
    sc = np.empty(n * k, np.float32)
    am = np.empty(n * k, np.int64)
    lib().orc_batch_f32(flat, offs, lens, n, strands.ravel(), strands.shape[1], k, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


def batch_f16(queries_h, strands_h, fused=False, n_threads=0):
    """Binary16 bit patterns in, binary16 score bits out."""
    strands = np.ascontiguousarray(np.atleast_2d(strands_h), np.uint16)
    flat, offs, lens = _pack(queries_h, np.uint16)
    n, k = len(queries_h), strands.shape[0]
    sc = np.empty(n * k, np.uint16)
    am = np.empty(n * k, np.int64)
    lib().orc_batch_f16(flat, offs, lens, n, strands.ravel(), strands.shape[1], k,
                        1 if fused else 0, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


def sdtw_f16_renorm(q_h, s_h, S, drift=0.01):
    """Renormalised binary16 form (bit patterns in) -> (float score, argmin)."""
    q = np.ascontiguousarray(q_h, np.uint16)
    s = np.ascontiguousarray(s_h, np.uint16)
    sc, am = C.c_float(), C.c_int64()
    lib().orc_sdtw_f16_renorm(q, q.size, s, s.size, S, drift, C.byref(sc), C.byref(am))
    return sc.value, am.value


def sdtw_f16_renorm_split(q_h, s_h, drift=0.01):
    """The automatic renormalised schedule for long reads: 512-row bands for
    the first 513..1536 rows, 1024-row bands for the rest."""
    q = np.ascontiguousarray(q_h, np.uint16)
    s = np.ascontiguousarray(s_h, np.uint16)
    sc, am = C.c_float(), C.c_int64()
    lib().orc_sdtw_f16_renorm_split(q, q.size, s, s.size, drift, C.byref(sc), C.byref(am))
    return sc.value, am.value


def sdtw_f16_renorm_row(q_h, s_h, S=8, split=False, drift=0.01):
    """(lowered last row as float32, binary32 offset) of the renormalised
    form; score = min(row) + offset at the first minimum."""
    q = np.ascontiguousarray(q_h, np.uint16)
    s = np.ascontiguousarray(s_h, np.uint16)
    row = np.empty(s.size, np.float32)
    off = C.c_float()
    lib().orc_sdtw_f16_renorm_row(q, q.size, s, s.size, S, 1 if split else 0, drift, row, C.byref(off))
    return row, np.float32(off.value)


def renorm_split_rows(Q):
    return int(lib().orc_renorm_split_rows(int(Q)))


def batch_f16_renorm(queries_h, strands_h, S, drift=0.01, n_threads=0):
    """S = 0: the automatic schedule (S = 8 up to 1536 rows, else the split)."""
    flat, offs, lens = _pack(queries_h, np.uint16)
    strands_h = np.ascontiguousarray(np.atleast_2d(strands_h), np.uint16)
    n, k = len(queries_h), strands_h.shape[0]
    sc = np.empty(n * k, np.float32)
    am = np.empty(n * k, np.int64)
    lib().orc_batch_f16_renorm(flat, offs, lens, n, strands_h.ravel(), strands_h.shape[1], k, S,
                               drift, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k)


_BASELINE = None
_V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def cpu_flags():
    """(CPU model name, set of flags) of this host, from /proc/cpuinfo; the
    name carries the CPU family and model numbers (a virtualised host may
    report only a brand string)."""
    model, flags, fam = "unknown", set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key = line.split(":", 1)[0].strip()
                if key == "model name" and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif key in ("cpu family", "model") and key not in fam:
                    fam[key] = line.split(":", 1)[1].strip()
                elif key == "flags" and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    if fam:
        model += " (family " + fam.get("cpu family", "?") + ", model " + fam.get("model", "?") + ")"
    return model, flags


def baseline_lib():
    """(library, ISA name) for the CPU baseline: the x86-64-v4 (AVX-512)
    build when this host has those instructions and it is built, else the
    portable AVX2 build (the checker's own library)."""
    global _BASELINE
    if _BASELINE is None:
        _, flags = cpu_flags()
        path = os.path.join(_HERE, "liboracle_v4.so")
        if all(f in flags for f in _V4_FLAGS) and os.path.exists(path):
            L = C.CDLL(path)
            L.orc_cpu_baseline_f32.restype = C.c_int
            L.orc_cpu_baseline_f32.argtypes = lib().orc_cpu_baseline_f32.argtypes
            _BASELINE = (L, "x86-64-v4 (AVX-512)")
        else:
            _BASELINE = (lib(), "x86-64-v3 (AVX2, FMA, F16C)")
    return _BASELINE


def cpu_baseline_f32(queries, strands, n_threads=0, portable=False):
    """FP32 CPU port of the recurrence (OpenMP over reads, SIMD over columns,
    column strips for long references) -> (scores, argmins, threads used).
    Runs the build for this host's best ISA unless portable=True."""
    strands = np.ascontiguousarray(np.atleast_2d(strands), np.float32)
    flat, offs, lens = _pack(queries, np.float32)
    n, k = len(queries), strands.shape[0]
    sc = np.empty(n * k, np.float32)
    am = np.empty(n * k, np.int64)
    L = lib() if portable else baseline_lib()[0]
    used = L.orc_cpu_baseline_f32(flat, offs, lens, n, strands.ravel(), strands.shape[1],
                                  k, sc, am, n_threads)
    return sc.reshape(n, k), am.reshape(n, k), used


def emulate_reference(ref, queries, query_len=2048, prefix_len=1024, segment=26,
                      warp=32, ref_len=48128, fp16=True, fix_q1=False, fix_q2=False,
                      n_threads=0, last_row=False):
    """Literal emulation of the reference kernel.  ref: fp16: (2, ref_len)
    fwd/rev values (already binary16-representable); fp32: (ref_len,).
    queries: (n_reads, query_len) normalised floats.  last_row=True also
    returns device_last_row as the kernel leaves it: (n_reads, ref_len, 2)
    for fp16, (n_reads, ref_len) for fp32, tiled index order, +inf where the
    kernel never writes."""
    g = EmuGeom(query_len, prefix_len, segment, warp, ref_len, 1 if fp16 else 0,
                1 if fix_q1 else 0, 1 if fix_q2 else 0)
    ref = np.ascontiguousarray(ref, np.float32).ravel()
    q = np.ascontiguousarray(queries, np.float32).ravel()
    n = q.size // query_len
    out = np.empty(n * (2 if fp16 else 1), np.float32)
    lr = np.empty(n * ref_len * (2 if fp16 else 1), np.float32) if last_row else None
    rc = lib().orc_emulate_reference(C.addressof(g), ref, q, n, out,
                                     None if lr is None else lr.ctypes.data, n_threads)
    if rc != 0:
        raise ValueError("unsupported emulator geometry")
    res = out.reshape(n, 2) if fp16 else out
    if last_row:
        return res, (lr.reshape(n, ref_len, 2) if fp16 else lr.reshape(n, ref_len))
    return res


# ------------------------------------------------------------ normalisation
def numpy_sum(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib().orc_numpy_sum(a, a.size)


def python_znorm(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib().orc_python_znorm(x, x.size, out)
    return out


def norm_meanstd(x, eps=1e-4):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    m, s = C.c_double(), C.c_double()
    lib().orc_norm_meanstd(x, x.size, eps, out, C.byref(m), C.byref(s))
    return out, m.value, s.value


def norm_medmad(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    m, s = C.c_double(), C.c_double()
    lib().orc_norm_medmad(x, x.size, out, C.byref(m), C.byref(s))
    return out, m.value, s.value


def squiggle_reference(fwd: str, rev: str, model_means, ref_len, k=6):
    mm = np.ascontiguousarray(model_means, np.float32)
    out = np.empty(2 * ref_len, np.float32)
    skipped = lib().orc_squiggle_reference(fwd.encode(), rev.encode(), len(fwd), k, mm,
                                           ref_len, out)
    return out.reshape(2, ref_len), skipped
