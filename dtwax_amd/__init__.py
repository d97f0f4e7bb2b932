"""dtwax_amd -- MI355X (gfx950) subsequence-DTW read classifier.

Python host-side mirror of the C ABI in include/dtwax.h (ctypes over the
in-tree ``libdtwax.so``).  Device memory is owned by the library's own HIP
calls (no torch).  If the library is missing the import of any entry point
raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DTWAX_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("DTWAX_LIB", os.path.join(_HERE, "libdtwax.so"))

F16, F16_FUSED, F16_RENORM, F32 = 0, 1, 2, 3
PRECISIONS = {"f16": F16, "f16_fused": F16_FUSED, "f16_renorm": F16_RENORM, "f32": F32}
NORM_NONE, NORM_MEANSTD, NORM_MEDMAD = 0, 1, 2

_lib = None


class DtwaxError(RuntimeError):
    pass


class ScoreOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("segment", C.c_int32), ("ref_chunk", C.c_int32),
                ("renorm_drift", C.c_float), ("compat_quirks", C.c_int32),
                ("moves", C.c_int32), ("waves", C.c_int32), ("packing", C.c_int32)]


RESULT_DTYPE = np.dtype([("score", np.float32, 2), ("argmin", np.int32, 2)])
TUNE_DTYPE = np.dtype([("segment", np.int32), ("ref_chunk", np.int32), ("ms", np.float32),
                       ("gcups", np.float32)])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DtwaxError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C "
                             "dtwax_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
        sig = {
            "dtwax_last_error": ([], C.c_char_p),
            "dtwax_version": ([], C.c_char_p),
            "dtwax_create": ([C.c_int, C.POINTER(vp)], C.c_int),
            "dtwax_destroy": ([vp], None),
            "dtwax_device_info": ([vp, C.POINTER(i32), C.POINTER(i32), C.c_char_p, i32], C.c_int),
            "dtwax_malloc": ([C.POINTER(vp), C.c_size_t], C.c_int),
            "dtwax_free": ([vp], C.c_int),
            "dtwax_memcpy_h2d": ([vp, vp, C.c_size_t, vp], C.c_int),
            "dtwax_memcpy_d2h": ([vp, vp, C.c_size_t, vp], C.c_int),
            "dtwax_stream_create": ([C.POINTER(vp)], C.c_int),
            "dtwax_stream_destroy": ([vp], C.c_int),
            "dtwax_stream_sync": ([vp], C.c_int),
            "dtwax_device_sync": ([], C.c_int),
            "dtwax_event_create": ([C.POINTER(vp)], C.c_int),
            "dtwax_event_destroy": ([vp], C.c_int),
            "dtwax_event_record": ([vp, vp], C.c_int),
            "dtwax_event_elapsed_ms": ([vp, vp, C.POINTER(f32)], C.c_int),
            "dtwax_reference_upload": ([vp, vp, i32, i64], C.c_int),
            "dtwax_squiggle_build": ([C.c_char_p, i64, i32, vp, i64, vp, C.POINTER(i64)], C.c_int),
            "dtwax_normalize": ([vp, vp, vp, vp, i32, i32, C.c_double, vp, vp, vp], C.c_int),
            "dtwax_score": ([vp, vp, vp, vp, i32, vp, vp, vp], C.c_int),
            "dtwax_plan_create": ([vp, vp, i32, vp, C.POINTER(vp)], C.c_int),
            "dtwax_plan_destroy": ([vp], None),
            "dtwax_plan_info": ([vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32), C.POINTER(i32),
                                 C.POINTER(i32), C.POINTER(i32)], C.c_int),
            "dtwax_score_plan": ([vp, vp, vp, vp, vp, vp], C.c_int),
            "dtwax_plan_set_timing": ([vp, i32], C.c_int),
            "dtwax_plan_packing": ([vp, C.POINTER(i32), C.POINTER(i32)], C.c_int),
            "dtwax_plan_kernel_ms": ([vp, C.POINTER(f32)], C.c_int),
            "dtwax_distances": ([vp, vp, vp, vp, i32, C.c_uint32, vp, vp, i32], C.c_int),
            "dtwax_distances_f32": ([vp, vp, vp, vp, i32, f32, vp, vp, i32], C.c_int),
            "dtwax_distances_noalloc": ([vp, vp, vp, vp, i32, C.c_uint32, vp, vp, i32], C.c_int),
            "dtwax_selftest_lanes": ([vp], C.c_int),
            "dtwax_autotune": ([vp, vp, vp, vp, i32, vp, i32, i64, C.c_char_p, vp, vp, i32,
                                C.POINTER(i32)], C.c_int),
            "dtwax_ru_create": ([vp, vp, C.POINTER(vp)], C.c_int),
            "dtwax_ru_destroy": ([vp], C.c_int),
            "dtwax_ru_info": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)],
                              C.c_int),
            "dtwax_ru_stats": ([vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64),
                                C.POINTER(C.c_double)], C.c_int),
            "dtwax_ru_process": ([vp, vp, i32, vp, vp], C.c_int),
            "dtwax_ru_release": ([vp, i32], C.c_int),
            "dtwax_ru_start": ([vp], C.c_int),
            "dtwax_ru_submit": ([vp, i32, i32, vp], C.c_int),
            "dtwax_ru_poll": ([vp, vp, i32, C.POINTER(i32), i32], C.c_int),
            "dtwax_ru_stop": ([vp], C.c_int),
            "dtwax_vbz_decode": ([vp, i64, vp, i32, vp, i64, C.POINTER(i64)], C.c_int),
            "dtwax_pod5_open": ([C.c_char_p, C.POINTER(vp)], C.c_int),
            "dtwax_pod5_close": ([vp], C.c_int),
            "dtwax_pod5_count": ([vp, C.POINTER(i64), C.POINTER(i64)], C.c_int),
            "dtwax_pod5_reads": ([vp, i64, i64, vp], C.c_int),
            "dtwax_pod5_signal": ([vp, vp, i64, i64, i64, vp, i64, vp, i32], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    """Names every C-ABI entry point in include/dtwax.h (checked by tests)."""
    return [
        "dtwax_last_error", "dtwax_version", "dtwax_create", "dtwax_destroy", "dtwax_device_info",
        "dtwax_malloc", "dtwax_free", "dtwax_memcpy_h2d", "dtwax_memcpy_d2h", "dtwax_stream_create",
        "dtwax_stream_destroy", "dtwax_stream_sync", "dtwax_device_sync", "dtwax_event_create",
        "dtwax_event_destroy", "dtwax_event_record", "dtwax_event_elapsed_ms",
        "dtwax_reference_upload", "dtwax_squiggle_build", "dtwax_normalize", "dtwax_score",
        "dtwax_plan_create", "dtwax_plan_destroy", "dtwax_plan_info", "dtwax_score_plan",
        "dtwax_plan_set_timing", "dtwax_plan_packing", "dtwax_plan_kernel_ms", "dtwax_distances", "dtwax_distances_f32",
        "dtwax_distances_noalloc", "dtwax_selftest_lanes",
        "dtwax_autotune", "dtwax_ru_create", "dtwax_ru_destroy", "dtwax_ru_info", "dtwax_ru_stats", "dtwax_ru_process",
        "dtwax_ru_release", "dtwax_ru_start", "dtwax_ru_submit", "dtwax_ru_poll", "dtwax_ru_stop",
        "dtwax_vbz_decode", "dtwax_pod5_open", "dtwax_pod5_close", "dtwax_pod5_count",
        "dtwax_pod5_reads", "dtwax_pod5_signal",
    ]


def _check(rc):
    if rc != 0:
        raise DtwaxError(f"dtwax error {rc}: {lib().dtwax_last_error().decode()}")


class DeviceArray:
    """A device allocation holding a numpy-shaped array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().dtwax_malloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _check(lib().dtwax_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, stream))
        return d

    def to_host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(lib().dtwax_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def free(self):
        if self.ptr:
            lib().dtwax_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pack_reads(reads):
    """List of 1-D sample arrays -> (flat float32, offsets int64, lengths int32)."""
    lens = np.array([len(r) for r in reads], np.int32)
    offs = np.zeros(len(reads), np.int64)
    if len(reads) > 1:
        offs[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    flat = (np.concatenate([np.asarray(r, np.float32) for r in reads]) if len(reads)
            else np.zeros(0, np.float32))
    return np.ascontiguousarray(flat, np.float32), offs, lens


def make_opts(precision=F16, segment=0, ref_chunk=0, renorm_drift=0.0, moves=2, waves=0, packing=0):
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    return ScoreOpts(int(precision), int(segment), int(ref_chunk), float(renorm_drift), 0,
                     int(moves), int(waves), int(packing))


class Plan:
    def __init__(self, ctx, lens, **kw):
        self.ctx = ctx
        lens = np.ascontiguousarray(lens, np.int32)
        self.n_reads = lens.size
        self.opts = make_opts(**kw)
        p = C.c_void_p()
        _check(lib().dtwax_plan_create(ctx.handle, lens.ctypes.data, lens.size,
                                       C.byref(self.opts), C.byref(p)))
        self.handle = p.value

    def info(self):
        cells, comp = C.c_int64(), C.c_int64()
        n_items, seg, chunk, grid = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().dtwax_plan_info(self.handle, C.byref(cells), C.byref(comp), C.byref(n_items),
                                     C.byref(seg), C.byref(chunk), C.byref(grid)))
        packs, packed = C.c_int32(), C.c_int32()
        _check(lib().dtwax_plan_packing(self.handle, C.byref(packs), C.byref(packed)))
        return dict(cells=cells.value, computed=comp.value, n_items=n_items.value,
                    segment=seg.value, ref_chunk=chunk.value, grid=grid.value,
                    packs=packs.value, packed_reads=packed.value)

    def run(self, d_queries, d_offsets, d_lens, d_out, stream=None):
        _check(lib().dtwax_score_plan(self.handle, d_queries.ptr, d_offsets.ptr, d_lens.ptr,
                                      d_out.ptr, stream))

    def set_timing(self, on=True):
        _check(lib().dtwax_plan_set_timing(self.handle, 1 if on else 0))

    def kernel_ms(self):
        ms = C.c_float()
        _check(lib().dtwax_plan_kernel_ms(self.handle, C.byref(ms)))
        return ms.value

    def close(self):
        if self.handle:
            lib().dtwax_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        p = C.c_void_p()
        _check(lib().dtwax_create(device, C.byref(p)))
        self.handle = p.value
        self.n_strands = 0
        self.n_cols = 0

    def close(self):
        if self.handle:
            lib().dtwax_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        n_cu, clk = C.c_int32(), C.c_int32()
        name = C.create_string_buffer(64)
        _check(lib().dtwax_device_info(self.handle, C.byref(n_cu), C.byref(clk), name, 64))
        return dict(n_cu=n_cu.value, clock_khz=clk.value, name=name.value.decode())

    def upload_reference(self, levels):
        levels = np.ascontiguousarray(np.atleast_2d(levels), np.float32)
        _check(lib().dtwax_reference_upload(self.handle, levels.ctypes.data, levels.shape[0],
                                            levels.shape[1]))
        self.n_strands, self.n_cols = levels.shape

    def normalize(self, reads, mode=NORM_MEDMAD, eps=0.0, stream=None):
        """reads: list of arrays -> list of normalised float32 arrays (via the GPU)."""
        flat, offs, lens = pack_reads(reads)
        d_raw, d_off, d_len = (DeviceArray.from_host(x) for x in (flat, offs, lens))
        d_out = DeviceArray(flat.shape, np.float32)
        d_st = DeviceArray((2 * max(1, len(reads)),), np.float64)
        _check(lib().dtwax_normalize(self.handle, d_raw.ptr, d_off.ptr, d_len.ptr, len(reads),
                                     mode, eps, d_out.ptr, d_st.ptr, stream))
        _check(lib().dtwax_stream_sync(stream))
        out = d_out.to_host()
        st = d_st.to_host()[: 2 * len(reads)].reshape(-1, 2)
        return [out[o:o + n] for o, n in zip(offs, lens)], st

    def score(self, reads, **kw):
        """reads: list of normalised float arrays -> (scores (n,2) f32, argmin (n,2) i32)."""
        flat, offs, lens = pack_reads(reads)
        d_q, d_off, d_len = (DeviceArray.from_host(x) for x in (flat, offs, lens))
        d_out = DeviceArray((max(1, len(reads)),), RESULT_DTYPE)
        opts = make_opts(**kw)
        _check(lib().dtwax_score(self.handle, d_q.ptr, d_off.ptr, d_len.ptr, len(reads),
                                 C.byref(opts), d_out.ptr, None))
        res = d_out.to_host()[: len(reads)]
        return res["score"].copy(), res["argmin"].copy()

    def plan(self, lens, **kw):
        return Plan(self, lens, **kw)

    def autotune(self, d_queries, d_offsets, lens, sample_reads=0, sample_cols=0, cache_path=None,
                 **kw):
        """Times the segment widths on a sample of the workload (device
        queries/offsets, host lengths) and takes the plan's chunk rule for the
        whole workload at the fastest width.  Returns
        (dict(segment=, ref_chunk=), table of timed candidates)."""
        lens = np.ascontiguousarray(lens, np.int32)
        base = make_opts(**kw)
        best = ScoreOpts()
        table = np.zeros(64, TUNE_DTYPE)
        n = C.c_int32()
        _check(lib().dtwax_autotune(self.handle, d_queries.ptr, d_offsets.ptr, lens.ctypes.data,
                                    len(lens), C.byref(base), sample_reads, sample_cols,
                                    None if cache_path is None else str(cache_path).encode(),
                                    C.byref(best), table.ctypes.data, len(table), C.byref(n)))
        return dict(segment=best.segment, ref_chunk=best.ref_chunk), table[: n.value]

    def distances(self, d_ref_tiled, d_query_h2, d_dists, num_entries, compat_quirks=1,
                  stream=None, d_last_row=None):
        """FullDTW::distances of the reference's FP16 build (DTW.hpp:285-295)
        on device buffers in its layouts: tiled half2 reference, half2 query
        words, half2 dists, optional last-row buffer (num_entries x 48128
        half2).  compat_quirks: 1 = the reference kernel's DP (seam
        deviations included), 0 = the clean recurrence, 2 = as 1 from the
        tile-by-tile kernel.  Asynchronous on `stream`."""
        _check(lib().dtwax_distances(self.handle, d_ref_tiled.ptr, d_query_h2.ptr, d_dists.ptr,
                                     num_entries, 0, stream,
                                     None if d_last_row is None else d_last_row.ptr, compat_quirks))

    def distances_f32(self, d_ref_tiled, d_query, d_dists, num_entries, compat_quirks=1,
                      stream=None, d_last_row=None):
        """The same for the reference's FP32 build: tiled float reference of
        the concatenated fwd | rev squiggle (96256 columns), float queries and
        dists, optional last-row buffer (num_entries x 96256 floats)."""
        _check(lib().dtwax_distances_f32(self.handle, d_ref_tiled.ptr, d_query.ptr, d_dists.ptr,
                                         num_entries, 0.0, stream,
                                         None if d_last_row is None else d_last_row.ptr,
                                         compat_quirks))


def squiggle_build(seq: str, model_means, ref_len=None, k=6):
    """Host-side squiggle build (load_reference.hpp:33-108 arithmetic) -> (2, ref_len)."""
    mm = np.ascontiguousarray(model_means, np.float32)
    if ref_len is None:
        ref_len = len(seq) - k + 1
    out = np.empty(2 * ref_len, np.float32)
    skipped = C.c_int64()
    _check(lib().dtwax_squiggle_build(seq.encode(), len(seq), k, mm.ctypes.data, ref_len,
                                      out.ctypes.data, C.byref(skipped)))
    return out.reshape(2, ref_len), skipped.value


def vbz_decode(chunk, n_max, cd_values=(0, 2, 1, 1)):
    """One VBZ-compressed FAST5 signal chunk (bytes) -> int16 samples.
    cd_values: the filter's (version, integer_size, delta_zigzag, zstd_level)."""
    buf = np.frombuffer(bytes(chunk), np.uint8)
    cd = np.ascontiguousarray(cd_values, np.uint32)
    out = np.empty(max(1, n_max), np.int16)
    n = C.c_int64()
    _check(lib().dtwax_vbz_decode(buf.ctypes.data, buf.size, cd.ctypes.data, cd.size,
                                  out.ctypes.data, n_max, C.byref(n)))
    return out[: n.value].copy()


POD5_READ_DTYPE = np.dtype([("n_samples", np.int64), ("start_sample", np.int64),
                            ("channel", np.int32), ("read_number", np.int32), ("n_rows", np.int32),
                            ("well", np.int32), ("calibration_offset", np.float32),
                            ("calibration_scale", np.float32), ("read_id", "S40")])
assert POD5_READ_DTYPE.itemsize == 80


class Pod5File:
    """A POD5 file read by the library's own parser (dtwax_pod5_*): the
    container footer, the Reads and Signal Arrow tables, and VBZ or plain
    int16 signal rows.  Host-side; no GPU needed."""

    def __init__(self, path):
        self._h = C.c_void_p()
        _check(lib().dtwax_pod5_open(os.fsencode(path), C.byref(self._h)))
        n, rows = C.c_int64(), C.c_int64()
        _check(lib().dtwax_pod5_count(self._h, C.byref(n), C.byref(rows)))
        self.n_reads, self.n_signal_rows = n.value, rows.value
        self._info = np.zeros(self.n_reads, POD5_READ_DTYPE)
        if self.n_reads:
            _check(lib().dtwax_pod5_reads(self._h, 0, self.n_reads, self._info.ctypes.data))

    def close(self):
        if self._h:
            lib().dtwax_pod5_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reads(self):
        """Per-read metadata (POD5_READ_DTYPE), table order."""
        return self._info.copy()

    def read_ids(self):
        return [r.decode() for r in self._info["read_id"]]

    def signal(self, reads=None, skip=0, take=None, threads=0):
        """Samples [skip, skip + take) of the listed reads (all by default).
        take=None: whole reads, returned as a list of int16 arrays; an int:
        a (n, take) int16 array and the per-read counts written."""
        idx = np.arange(self.n_reads, dtype=np.int64) if reads is None else \
            np.ascontiguousarray(reads, np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= self.n_reads):
            raise DtwaxError("read index outside the file")
        n_out = np.zeros(max(1, idx.size), np.int64)
        if take is None:
            have = np.maximum(0, self._info["n_samples"][idx] - skip) if idx.size else idx
            dst = np.empty(max(1, int(have.sum())), np.int16)
            _check(lib().dtwax_pod5_signal(self._h, idx.ctypes.data, idx.size, skip, -1,
                                           dst.ctypes.data, int(have.sum()), n_out.ctypes.data, threads))
            ends = np.cumsum(have)
            return [dst[e - h:e].copy() for e, h in zip(ends, have)]
        dst = np.zeros((max(1, idx.size), max(1, take)), np.int16)
        _check(lib().dtwax_pod5_signal(self._h, idx.ctypes.data, idx.size, skip, take,
                                       dst.ctypes.data, idx.size * take, n_out.ctypes.data, threads))
        return dst[: idx.size, :take], n_out[: idx.size]


POD5_RESULT_DTYPE = np.dtype([("read_id", "U36"), ("score", np.float32, 2), ("argmin", np.int32, 2),
                              ("target", np.bool_)])


def classify_pod5(ctx, path, skip=1952, take=2048, norm=NORM_MEANSTD, eps=1e-4, threshold=None,
                  threads=0, **score_kw):
    """Scores a POD5 file's reads as dtwax_main scores its input: samples
    [skip, skip + take) of every read that has that many (shorter reads are
    left out), normalised on the GPU (mean/std with eps by default, as
    normalizer.cu:86-100), against both strands of the reference uploaded to
    ctx.  Returns a POD5_RESULT_DTYPE array in file order; `target` is
    min(fwd, rev) < threshold when a threshold is given."""
    with Pod5File(path) as f:
        info = f.reads()
        sel = np.nonzero(info["n_samples"] >= skip + take)[0]
        out = np.zeros(sel.size, POD5_RESULT_DTYPE)
        if not sel.size:
            return out
        win, _ = f.signal(reads=sel, skip=skip, take=take, threads=threads)
    normed, _ = ctx.normalize(list(win.astype(np.float32)), mode=norm, eps=eps)
    sc, am = ctx.score(normed, **score_kw)
    out["read_id"] = [info["read_id"][i].decode() for i in sel]
    out["score"], out["argmin"] = sc, am
    if threshold is not None:
        out["target"] = sc.min(axis=1) < threshold
    return out


def read_fasta(path):
    """Sequence of a single-entry FASTA, as load_reference::ref_loader reads
    it (load_reference.hpp:164-212): lines after a '>' header are joined; an
    entry whose sequence lines contain a space is dropped; with several
    entries the last one is used.  Raises ValueError if there is none."""
    name, content, seq = None, [], None
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n").rstrip("\r")
            if not line or line[0] == ">":
                if name:
                    seq = "".join(content)
                name = line[1:] if line else None
                content = []
            elif name:
                if " " in line:
                    name, content = None, []
                else:
                    content.append(line)
    if name:
        seq = "".join(content)
    if seq is None:
        raise ValueError(f"no valid FASTA entry in {path}")
    return seq


def read_kmer_model(path):
    """k-mer model "KMER mean sd" -> means[4**k] float32 indexed by 2-bit code
    (A=0 C=1 G=2 T=3, first base most significant); load_reference.hpp:131-161,
    the first line of a k-mer wins."""
    m = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) < 3:
                continue
            try:
                mean, _ = float(p[1]), float(p[2])
            except ValueError:
                continue
            m.setdefault(p[0], mean)
    if not m:
        raise ValueError(f"empty k-mer model {path}")
    k = len(min(m))
    means = np.zeros(4 ** k, np.float32)
    codes = {"A": 0, "C": 1, "G": 2, "T": 3}
    for kmer, v in m.items():
        if len(kmer) != k or any(ch not in codes for ch in kmer):
            continue
        c = 0
        for ch in kmer:
            c = c * 4 + codes[ch]
        means[c] = np.float32(v)
    return means


def selftest_lanes():
    out = np.empty(128, np.int32)
    _check(lib().dtwax_selftest_lanes(out.ctypes.data))
    return out[:64], out[64:]


# ---- Read Until serving ----------------------------------------------------
RU_CONTINUE, RU_NONTARGET, RU_TARGET, RU_NO_SLOT, RU_STALE = 0, 1, 2, 3, 4
RU_ACTIONS = {RU_CONTINUE: "continue", RU_NONTARGET: "nontarget", RU_TARGET: "target",
              RU_NO_SLOT: "no_slot", RU_STALE: "stale"}


class RuOpts(C.Structure):
    _fields_ = [("max_reads", C.c_int32), ("chunk_len", C.c_int32), ("max_chunks", C.c_int32),
                ("precision", C.c_int32), ("segment", C.c_int32), ("norm", C.c_int32),
                ("eps", C.c_double), ("threshold", C.c_float), ("min_ref_chunk", C.c_int32),
                ("hbm_budget", C.c_int64), ("threshold_per_row", C.c_float),
                ("n_streams", C.c_int32), ("renorm_drift", C.c_float), ("reserved", C.c_int32)]


RU_CHUNK_DTYPE = np.dtype([("channel", np.int32), ("number", np.int32), ("offset", np.int64)])
RU_DECISION_DTYPE = np.dtype([("channel", np.int32), ("number", np.int32), ("action", np.int32),
                              ("chunks", np.int32), ("score", np.float32, 2),
                              ("argmin", np.int32, 2), ("latency_us", np.float64)])


class ReadUntil:
    """Stateful chunked classifier (dtwax_ru_* in include/dtwax.h) on a
    context's current reference."""

    def __init__(self, ctx, max_reads=512, chunk_len=1600, max_chunks=4, precision=F16,
                 segment=0, norm=NORM_MEDMAD, eps=0.0, threshold=0.0, min_ref_chunk=0,
                 hbm_budget=0, threshold_per_row=0.0, n_streams=0, renorm_drift=0.0):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        o = RuOpts(max_reads=max_reads, chunk_len=chunk_len, max_chunks=max_chunks,
                   precision=precision, segment=segment, norm=norm, eps=eps,
                   threshold=threshold, min_ref_chunk=min_ref_chunk, hbm_budget=hbm_budget,
                   threshold_per_row=threshold_per_row, n_streams=n_streams,
                   renorm_drift=renorm_drift)
        p = C.c_void_p()
        _check(lib().dtwax_ru_create(ctx.handle, C.byref(o), C.byref(p)))
        self.handle = p.value
        self.ctx = ctx  # the engine reads the context's reference
        self.chunk_len = chunk_len

    def info(self):
        sl, act, seg, nb = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        _check(lib().dtwax_ru_info(self.handle, C.byref(sl), C.byref(act), C.byref(seg), C.byref(nb)))
        return dict(slots=sl.value, active=act.value, segment=seg.value, state_bytes=nb.value)

    def stats(self):
        b, ch, ce, ms = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().dtwax_ru_stats(self.handle, C.byref(b), C.byref(ch), C.byref(ce), C.byref(ms)))
        return dict(batches=b.value, chunks=ch.value, cells=ce.value, gpu_ms=ms.value)

    def process(self, channels, numbers, samples):
        """One batch: samples (n, chunk_len) float32, one row per channel.
        Returns a RU_DECISION_DTYPE array."""
        samples = np.ascontiguousarray(samples, np.float32).reshape(-1, self.chunk_len)
        n = samples.shape[0]
        ch = np.zeros(n, RU_CHUNK_DTYPE)
        ch["channel"] = channels
        ch["number"] = numbers
        ch["offset"] = np.arange(n, dtype=np.int64) * self.chunk_len
        out = np.zeros(n, RU_DECISION_DTYPE)
        _check(lib().dtwax_ru_process(self.handle, ch.ctypes.data, n, samples.ctypes.data,
                                      out.ctypes.data))
        return out

    def release(self, channel):
        _check(lib().dtwax_ru_release(self.handle, channel))

    def start(self):
        _check(lib().dtwax_ru_start(self.handle))

    def submit(self, channel, number, samples):
        samples = np.ascontiguousarray(samples, np.float32)
        if samples.size != self.chunk_len:
            raise ValueError(f"a chunk holds {self.chunk_len} samples")
        _check(lib().dtwax_ru_submit(self.handle, channel, number, samples.ctypes.data))

    def poll(self, cap=1024, timeout_ms=0):
        out = np.zeros(cap, RU_DECISION_DTYPE)
        n = C.c_int32()
        _check(lib().dtwax_ru_poll(self.handle, out.ctypes.data, cap, C.byref(n), timeout_ms))
        return out[: n.value]

    def stop(self):
        _check(lib().dtwax_ru_stop(self.handle))

    def close(self):
        if self.handle:
            lib().dtwax_ru_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
