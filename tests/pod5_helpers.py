"""Test-side POD5 writer (test infrastructure).

Neither the pod5 package nor a POD5 file is available in this image, so the
reader (dtwax_pod5_*) is checked against files written here from the
published layout:

  signature "\\x8BPOD\\r\\n\\x1A\\n" | section marker (16 bytes)
  | Arrow IPC file | zero padding to 8 | section marker | ... (Signal, Run Info, Reads)
  | "FOOTER\\0\\0" | footer flatbuffer | int64 LE footer length
  | section marker | signature

The embedded tables are written by pyarrow (an independent Arrow
implementation, so the reader's Arrow parsing is checked against real Arrow
IPC files); the footer flatbuffer by the small builder below, which is
itself checked by having pyarrow parse an Arrow schema message built with it
(tests/test_pod5.py).  VBZ rows are svb16 blocks (1-bit keys, zig-zag 16-bit
differences) compressed by the system zstd.
"""
import io
import struct
import uuid as _uuid

import numpy as np
import pyarrow as pa
import pyarrow.ipc

from vbz_helpers import zstd_compress

SIGNATURE = b"\x8bPOD\r\n\x1a\n"
FOOTER_MAGIC = b"FOOTER\x00\x00"
READS_TABLE, SIGNAL_TABLE, RUN_INFO_TABLE = 0, 1, 4


# ---- flatbuffers (write side) ----------------------------------------------
class FbTable:
    """slots: list indexed by field id; each None or (kind, value).
    kinds: u8 bool i16 i32 i64 (scalars), str, table (FbTable),
    tables ([FbTable]), structs ((bytes, align, count))."""

    def __init__(self, *slots):
        self.slots = list(slots)


_SCALAR = {"u8": ("<B", 1), "bool": ("<B", 1), "i16": ("<h", 2), "i32": ("<i", 4), "i64": ("<q", 8)}


def _align(buf, a):
    while len(buf) % a:
        buf.append(0)


def _emit_table(buf, t):
    """Appends t (vtable, then the table), then its children after it (so
    every uoffset points forward); returns the table position."""
    n = len(t.slots)
    _align(buf, 2)
    vt_pos = len(buf)
    vt_size = 4 + 2 * n
    # lay the table out: soffset, then the fields, each at an absolute
    # position aligned to its size
    tpos = vt_pos + vt_size
    tpos += (-tpos) % 4
    cur = tpos + 4
    place = []
    for s in t.slots:
        if s is None:
            place.append(None)
            continue
        kind = s[0]
        size = _SCALAR[kind][1] if kind in _SCALAR else 4
        cur += (-cur) % size
        place.append(cur)
        cur += size
    tsize = cur - tpos
    offs = [0 if p is None else p - tpos for p in place]
    buf.extend(b"\x00" * (vt_pos - len(buf)))
    buf.extend(struct.pack("<HH", vt_size, tsize) + b"".join(struct.pack("<H", o) for o in offs))
    buf.extend(b"\x00" * (tpos - len(buf)))
    buf.extend(b"\x00" * tsize)
    struct.pack_into("<i", buf, tpos, tpos - vt_pos)
    pending = []
    for s, p in zip(t.slots, place):
        if s is None:
            continue
        kind, val = s
        if kind in _SCALAR:
            struct.pack_into(_SCALAR[kind][0], buf, p, int(val))
        else:
            pending.append((kind, val, p))
    for kind, val, p in pending:
        target = _emit_child(buf, kind, val)
        struct.pack_into("<I", buf, p, target - p)
    return tpos


def _emit_child(buf, kind, val):
    if kind == "str":
        b = val.encode() if isinstance(val, str) else bytes(val)
        _align(buf, 4)
        pos = len(buf)
        buf.extend(struct.pack("<I", len(b)) + b + b"\x00")
        return pos
    if kind == "table":
        return _emit_table(buf, val)
    if kind == "structs":
        data, align, count = val
        _align(buf, 4)
        while (len(buf) + 4) % align:
            buf.extend(b"\x00" * 4)
        pos = len(buf)
        buf.extend(struct.pack("<I", count) + bytes(data))
        return pos
    if kind == "tables":
        _align(buf, 4)
        pos = len(buf)
        buf.extend(struct.pack("<I", len(val)) + b"\x00" * (4 * len(val)))
        for i, t in enumerate(val):
            slot = pos + 4 + 4 * i
            tp = _emit_table(buf, t)
            struct.pack_into("<I", buf, slot, tp - slot)
        return pos
    raise ValueError(kind)


def fb_build(root):
    buf = bytearray(b"\x00" * 8)  # root uoffset, then 4 bytes so the table may sit 8-aligned
    tpos = _emit_table(buf, root)
    struct.pack_into("<I", buf, 0, tpos)
    _align(buf, 8)
    return bytes(buf)


def arrow_schema_message(fields, metadata):
    """An Arrow IPC schema message built with fb_build (to check the builder
    against pyarrow's parser).  fields: [(name, type_id, type FbTable,
    children, dictionary FbTable or None, {key: value})]."""
    def kvs(d):
        return ("tables", [FbTable(("str", k), ("str", v)) for k, v in d.items()])

    def field(f):
        name, tid, ty, kids, dic, md = f
        return FbTable(("str", name), ("bool", 1), ("u8", tid), ("table", ty),
                       ("table", dic) if dic is not None else None,
                       ("tables", [field(k) for k in kids]), kvs(md) if md else None)

    schema = FbTable(("i16", 0), ("tables", [field(f) for f in fields]), kvs(metadata))
    msg = FbTable(("i16", 4), ("u8", 1), ("table", schema), ("i64", 0))
    fb = fb_build(msg)
    return b"\xff\xff\xff\xff" + struct.pack("<i", len(fb)) + fb


# ---- svb16 + zstd (POD5's VBZ signal rows) ---------------------------------
def svb16_encode(samples):
    x = np.asarray(samples, np.int16).view(np.uint16).astype(np.uint32)
    prev = np.concatenate([[0], x[:-1]]).astype(np.uint32)
    d = (x - prev) & 0xFFFF
    zz = ((d << 1) & 0xFFFF) ^ np.where(d & 0x8000, 0xFFFF, 0)
    two = zz > 0xFF
    keys = np.packbits(two.astype(np.uint8), bitorder="little") if x.size else np.zeros(0, np.uint8)
    lens = 1 + two.astype(np.int64)
    pos = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    data = np.zeros(int(lens.sum()), np.uint8)
    data[pos] = zz & 0xFF
    data[pos[two] + 1] = zz[two] >> 8
    return keys.tobytes() + data.tobytes()


def vbz_row(samples, level=1):
    return zstd_compress(svb16_encode(samples), level)


# ---- POD5 -------------------------------------------------------------------
class UuidType(pa.ExtensionType):
    def __init__(self):
        super().__init__(pa.binary(16), "minknow.uuid")

    def __arrow_ext_serialize__(self):
        return b""

    @classmethod
    def __arrow_ext_deserialize__(cls, storage_type, serialized):
        return cls()


class VbzType(pa.ExtensionType):
    def __init__(self):
        super().__init__(pa.large_binary(), "minknow.vbz")

    def __arrow_ext_serialize__(self):
        return b""

    @classmethod
    def __arrow_ext_deserialize__(cls, storage_type, serialized):
        return cls()


def _ipc(schema, batches, compression):
    sink = io.BytesIO()
    opts = pa.ipc.IpcWriteOptions(compression=compression)
    with pa.ipc.new_file(sink, schema, options=opts) as w:
        for b in batches:
            w.write_batch(b)
    return sink.getvalue()


def _batches(schema, columns, n, per):
    per = max(1, per)
    out = []
    for a in range(0, max(n, 1), per):
        z = min(n, a + per)
        out.append(pa.record_batch([c[a:z] for c in columns], schema=schema))
    return out


def make_read(signal, i=0, **kw):
    r = dict(read_id=str(_uuid.UUID(int=(0x1234 << 100) + 7919 * i + 1)), signal=np.asarray(signal, np.int16),
             channel=1 + i % 512, well=1 + i % 4, read_number=100 + i, start=1000 * i,
             calibration_offset=-240.0 + i, calibration_scale=0.17 + 0.001 * i)
    r.update(kw)
    return r


def write_pod5(path, reads, vbz=True, row_samples=102400, signal_batch=1000, read_batch=1000,
               shuffle_rows=None, ipc_compression=None, footer_pad=0, file_id=None,
               table_file_id=None, marker=None, bad_row=None, num_samples_delta=None,
               samples_delta=None, omit=()):
    """reads: [make_read(...)].  Options that break the file on purpose:
    bad_row=(read, index) names a row past the table; num_samples_delta=(read,
    d) and samples_delta=(row, d) offset a count; omit drops columns."""
    file_id = file_id or str(_uuid.UUID(int=0xABCDEF))
    marker = marker or _uuid.UUID(int=0x5EC7104).bytes
    meta = {"MINKNOW:pod5_version": "0.3.10", "MINKNOW:software": "dtwax test writer",
            "MINKNOW:file_identifier": table_file_id or file_id}

    # signal rows, read by read, then optionally permuted
    row_read, row_sig = [], []
    read_rows = []
    for ri, r in enumerate(reads):
        sig = r["signal"]
        rows = []
        for a in range(0, sig.size, row_samples):
            rows.append(len(row_sig))
            row_read.append(ri)
            row_sig.append(sig[a:a + row_samples])
        read_rows.append(rows)
    perm = np.arange(len(row_sig))
    if shuffle_rows is not None:
        perm = np.random.default_rng(shuffle_rows).permutation(len(row_sig))
    where = np.empty(len(row_sig), np.int64)
    where[perm] = np.arange(len(row_sig))  # old row -> new position
    sig_ids = [_uuid.UUID(reads[row_read[o]]["read_id"]).bytes for o in perm]
    counts = [row_sig[o].size for o in perm]
    if samples_delta is not None:
        counts[samples_delta[0]] += samples_delta[1]
    if vbz:
        sig_col = pa.ExtensionArray.from_storage(
            VbzType(), pa.array([vbz_row(row_sig[o]) for o in perm], pa.large_binary()))
    else:
        offs = np.concatenate([[0], np.cumsum([row_sig[o].size for o in perm])]).astype(np.int64)
        vals = np.concatenate([row_sig[o] for o in perm]) if perm.size else np.zeros(0, np.int16)
        sig_col = pa.LargeListArray.from_arrays(pa.array(offs, pa.int64()), pa.array(vals, pa.int16()))
    sig_schema = pa.schema([("read_id", UuidType()), ("signal", sig_col.type), ("samples", pa.uint32())],
                           metadata=meta)
    sig_cols = [pa.ExtensionArray.from_storage(UuidType(), pa.array(sig_ids, pa.binary(16))), sig_col,
                pa.array(counts, pa.uint32())]
    signal_ipc = _ipc(sig_schema, _batches(sig_schema, sig_cols, len(counts), signal_batch), ipc_compression)

    run_schema = pa.schema([("acquisition_id", pa.utf8()), ("sample_rate", pa.uint16()),
                            ("adc_min", pa.int16()), ("adc_max", pa.int16())], metadata=meta)
    run_ipc = _ipc(run_schema, [pa.record_batch([pa.array(["acq0"]), pa.array([4000], pa.uint16()),
                                                 pa.array([-4096], pa.int16()), pa.array([4095], pa.int16())],
                                                schema=run_schema)], ipc_compression)

    n = len(reads)
    row_lists = [[int(where[o]) for o in rows] for rows in read_rows]
    if bad_row is not None:
        row_lists[bad_row[0]][bad_row[1]] = len(counts) + 3
    nsamp = [int(r["signal"].size) for r in reads]
    if num_samples_delta is not None:
        nsamp[num_samples_delta[0]] += num_samples_delta[1]
    pore = pa.DictionaryArray.from_arrays(pa.array([0] * n, pa.int16()), pa.array(["not_set", "r9"]))
    endr = pa.DictionaryArray.from_arrays(pa.array([i % 3 for i in range(n)], pa.int16()),
                                          pa.array(["unknown", "signal_positive", "data_service_unblock"]))
    runinfo = pa.DictionaryArray.from_arrays(pa.array([0] * n, pa.int16()), pa.array(["acq0"]))
    cols = [
        ("read_id", pa.ExtensionArray.from_storage(UuidType(), pa.array(
            [_uuid.UUID(r["read_id"]).bytes for r in reads], pa.binary(16)))),
        ("signal", pa.array(row_lists, pa.list_(pa.uint64()))),
        ("read_number", pa.array([r["read_number"] for r in reads], pa.uint32())),
        ("start", pa.array([r["start"] for r in reads], pa.uint64())),
        ("median_before", pa.array([200.0] * n, pa.float32())),
        ("num_minknow_events", pa.array([0] * n, pa.uint64())),
        ("tracked_scaling_scale", pa.array([np.nan] * n, pa.float32())),
        ("tracked_scaling_shift", pa.array([np.nan] * n, pa.float32())),
        ("predicted_scaling_scale", pa.array([np.nan] * n, pa.float32())),
        ("predicted_scaling_shift", pa.array([np.nan] * n, pa.float32())),
        ("num_reads_since_mux_change", pa.array([0] * n, pa.uint32())),
        ("time_since_mux_change", pa.array([0.0] * n, pa.float32())),
        ("channel", pa.array([r["channel"] for r in reads], pa.uint16())),
        ("well", pa.array([r["well"] for r in reads], pa.uint8())),
        ("pore_type", pore),
        ("calibration_offset", pa.array([r["calibration_offset"] for r in reads], pa.float32())),
        ("calibration_scale", pa.array([r["calibration_scale"] for r in reads], pa.float32())),
        ("end_reason", endr),
        ("end_reason_forced", pa.array([False] * n, pa.bool_())),
        ("run_info", runinfo),
        ("num_samples", pa.array(nsamp, pa.uint64())),
    ]
    cols = [c for c in cols if c[0] not in omit]
    reads_schema = pa.schema([(k, v.type) for k, v in cols], metadata=meta)
    reads_ipc = _ipc(reads_schema, _batches(reads_schema, [v for _, v in cols], n, read_batch),
                     ipc_compression)

    out = bytearray(SIGNATURE + marker)
    contents = []
    for ct, blob in ((SIGNAL_TABLE, signal_ipc), (RUN_INFO_TABLE, run_ipc), (READS_TABLE, reads_ipc)):
        contents.append(FbTable(("i64", len(out)), ("i64", len(blob)), ("i16", 0), ("i16", ct)))
        out += blob
        _align(out, 8)
        out += marker
    footer = fb_build(FbTable(("str", file_id), ("str", "dtwax test writer"), ("str", "0.3.10"),
                              ("tables", contents)))
    out += FOOTER_MAGIC + footer + b"\x00" * footer_pad
    out += struct.pack("<q", len(footer)) + marker + SIGNATURE
    with open(path, "wb") as f:
        f.write(bytes(out))
    return dict(signal=signal_ipc, run_info=run_ipc, reads=reads_ipc, footer=footer)
