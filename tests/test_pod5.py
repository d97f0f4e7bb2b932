"""POD5 input (dtwax_pod5_*) on the CPU, and through the CLI.

Parity unpinned against the pod5 library itself: neither it nor a POD5 file
is available in this image (the reference reads FAST5 only,
fast5/nanopolish_fast5_io.cpp).  What is pinned: the embedded tables are
real Arrow IPC files written by pyarrow and the reader's view of them is
compared with pyarrow's own; the footer flatbuffer builder of the test
writer is checked by pyarrow's flatbuffer parser; VBZ rows round-trip an
independent numpy svb16 encoder; and a FAST5 folder and a POD5 holding the
same reads give the CLI the same reads.
"""
import os
import struct
import subprocess
import sys
import uuid

import numpy as np
import pyarrow as pa
import pyarrow.ipc
import pytest

import dtwax_amd as D
from helpers import GOLDEN
from pod5_helpers import (FbTable, arrow_schema_message, make_read, svb16_encode, vbz_row,
                          write_pod5)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dtwax_amd", "bin", "dtwax_main")


def _reads(lens, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(lens):
        x = np.cumsum(rng.integers(-300, 301, n)) + rng.integers(-2000, 2000)
        x = np.clip(x, -32768, 32767).astype(np.int16)
        if n > 10:
            x[3], x[4], x[5] = 32767, -32768, 32767  # full-range jumps (16-bit wrap of the difference)
        out.append(make_read(x, i))
    return out


def test_flatbuffer_builder_pinned_by_pyarrow():
    """The writer's flatbuffer builder, used for the POD5 footer, produces
    messages that pyarrow's parser (with flatbuffers verification) reads back
    field by field."""
    Int = lambda bits, signed: FbTable(("i32", bits), ("bool", int(signed)))
    fields = [
        ("read_id", 15, FbTable(("i32", 16)), [], None, {"ARROW:extension:name": "minknow.uuid"}),
        ("signal", 21, FbTable(), [("item", 2, Int(16, True), [], None, {})], None, {}),
        ("samples", 2, Int(32, False), [], None, {}),
        ("pore_type", 5, FbTable(), [], FbTable(("i64", 0), ("table", Int(16, True))), {}),
        ("cal", 3, FbTable(("i16", 2)), [], None, {}),
        ("start", 2, Int(64, False), [], None, {}),
    ]
    s = pa.ipc.read_schema(pa.py_buffer(arrow_schema_message(fields, {"MINKNOW:file_identifier": "xyz"})))
    assert s.names == ["read_id", "signal", "samples", "pore_type", "cal", "start"]
    assert s.field("read_id").type == pa.binary(16)
    assert s.field("read_id").metadata == {b"ARROW:extension:name": b"minknow.uuid"}
    assert s.field("signal").type == pa.large_list(pa.int16())
    assert s.field("samples").type == pa.uint32()
    assert s.field("pore_type").type == pa.dictionary(pa.int16(), pa.utf8())
    assert s.field("cal").type == pa.float64()
    assert s.field("start").type == pa.uint64()
    assert s.metadata == {b"MINKNOW:file_identifier": b"xyz"}


def test_svb16_hand_encoded():
    # samples 1 -2 300 300 -> differences 1 -3 302 0 -> zig-zag 2 5 604 0
    # keys: only the third needs 2 bytes -> 0b0100; data 02 05 5c 02 00
    assert svb16_encode([1, -2, 300, 300]) == bytes([0x04, 0x02, 0x05, 0x5C, 0x02, 0x00])
    # 32767 -> -32768: the 16-bit difference wraps to +1 (zig-zag 2)
    assert svb16_encode([32767, -32768]) == bytes([0x01, 0xFE, 0xFF, 0x02])


@pytest.mark.parametrize("vbz", [True, False])
@pytest.mark.parametrize("compression", [None, "zstd"])
@pytest.mark.parametrize("shuffle", [None, 7])
def test_round_trip(tmp_path, vbz, compression, shuffle):
    reads = _reads([5000, 0, 250_000, 1, 12_345, 4096, 4097])
    p = tmp_path / "a.pod5"
    write_pod5(p, reads, vbz=vbz, row_samples=4096, signal_batch=5, read_batch=3,
               shuffle_rows=shuffle, ipc_compression=compression)
    with D.Pod5File(p) as f:
        assert f.n_reads == len(reads)
        assert f.n_signal_rows == sum(-(-r["signal"].size // 4096) for r in reads)
        info = f.reads()
        assert f.read_ids() == [r["read_id"] for r in reads]
        for k in ("channel", "well", "read_number"):
            assert list(info[k]) == [r[k] for r in reads]
        assert list(info["start_sample"]) == [r["start"] for r in reads]
        assert np.array_equal(info["calibration_offset"], np.float32([r["calibration_offset"] for r in reads]))
        assert np.array_equal(info["calibration_scale"], np.float32([r["calibration_scale"] for r in reads]))
        assert list(info["n_samples"]) == [r["signal"].size for r in reads]
        assert list(info["n_rows"]) == [-(-r["signal"].size // 4096) for r in reads]
        for got, r in zip(f.signal(), reads):
            assert np.array_equal(got, r["signal"])


def _embedded(path):
    """(offset, length, content type) of each embedded table, read with the
    struct module from the footer the test writer laid out."""
    b = open(path, "rb").read()
    flen = struct.unpack_from("<q", b, len(b) - 32)[0]
    fb = b[len(b) - 32 - flen: len(b) - 32]
    root = struct.unpack_from("<I", fb, 0)[0]
    vt = root - struct.unpack_from("<i", fb, root)[0]
    o_contents = struct.unpack_from("<H", fb, vt + 4 + 2 * 3)[0]
    vpos = root + o_contents + struct.unpack_from("<I", fb, root + o_contents)[0]
    out = []
    for i in range(struct.unpack_from("<I", fb, vpos)[0]):
        slot = vpos + 4 + 4 * i
        t = slot + struct.unpack_from("<I", fb, slot)[0]
        tvt = t - struct.unpack_from("<i", fb, t)[0]
        offs = [struct.unpack_from("<H", fb, tvt + 4 + 2 * k)[0] for k in range(4)]
        off, ln = (struct.unpack_from("<q", fb, t + offs[k])[0] for k in (0, 1))
        ct = struct.unpack_from("<h", fb, t + offs[3])[0]
        out.append((off, ln, ct))
    return b, out


@pytest.mark.parametrize("compression", [None, "zstd"])
def test_reader_agrees_with_pyarrow(tmp_path, compression):
    """The reader's view of the Arrow tables equals pyarrow's: read ids, row
    lists, per-row sample counts and (plain rows) the int16 values."""
    reads = _reads([9000, 3, 20_000, 0, 777], seed=3)
    p = tmp_path / "b.pod5"
    write_pod5(p, reads, vbz=False, row_samples=2500, signal_batch=4, read_batch=2, shuffle_rows=11,
               ipc_compression=compression)
    b, emb = _embedded(p)
    tables = {ct: pa.ipc.open_file(pa.py_buffer(b[off:off + ln])).read_all() for off, ln, ct in emb}
    assert set(tables) == {0, 1, 4}
    rt, st = tables[0], tables[1]
    col = rt.column("read_id").combine_chunks()
    col = getattr(col, "storage", col)  # fixed_size_binary(16) under the minknow.uuid name
    ids = [str(uuid.UUID(bytes=x)) for x in col.to_pylist()]
    rows = rt.column("signal").to_pylist()
    sig_rows = st.column("signal").to_pylist()
    with D.Pod5File(p) as f:
        assert f.read_ids() == ids
        assert list(f.reads()["n_rows"]) == [len(r) for r in rows]
        got = f.signal()
    for g, rl in zip(got, rows):
        want = np.concatenate([np.asarray(sig_rows[k], np.int16) for k in rl]) if rl else np.zeros(0, np.int16)
        assert np.array_equal(g, want)


def test_windows_and_threads(tmp_path):
    reads = _reads([10_000, 4096, 100, 0, 30_001], seed=5)
    p = tmp_path / "w.pod5"
    write_pod5(p, reads, vbz=True, row_samples=4096, signal_batch=3, shuffle_rows=2)
    with D.Pod5File(p) as f:
        for skip, take in [(0, 4096), (4095, 2), (4096, 4096), (1952, 2048), (9999, 10), (50_000, 5),
                           (0, 0), (8191, 8193)]:
            w1, n1 = f.signal(skip=skip, take=take, threads=1)
            w8, n8 = f.signal(skip=skip, take=take, threads=8)
            assert np.array_equal(w1, w8) and np.array_equal(n1, n8)
            for j, r in enumerate(reads):
                want = r["signal"][skip:skip + take]
                assert n1[j] == want.size
                assert np.array_equal(w1[j, :want.size], want)
        for skip in (0, 100, 5000):
            for g, r in zip(f.signal(skip=skip), reads):
                assert np.array_equal(g, r["signal"][skip:])
        sub = f.signal(reads=[4, 0, 4], skip=3, take=9)
        assert np.array_equal(sub[0][0], reads[4]["signal"][3:12])
        assert np.array_equal(sub[0][1], reads[0]["signal"][3:12])
        with pytest.raises(D.DtwaxError, match="outside"):
            f.signal(reads=[5])
        with pytest.raises(D.DtwaxError, match="negative"):
            f.signal(skip=-1, take=3)


def test_optional_columns_absent(tmp_path):
    reads = _reads([300, 5000])
    p = tmp_path / "o.pod5"
    write_pod5(p, reads, omit=("channel", "well", "read_number", "start", "calibration_offset",
                               "calibration_scale", "num_samples", "pore_type", "end_reason"))
    with D.Pod5File(p) as f:
        info = f.reads()
        assert list(info["channel"]) == [-1, -1] and list(info["well"]) == [-1, -1]
        assert list(info["read_number"]) == [-1, -1] and list(info["start_sample"]) == [0, 0]
        assert np.all(np.isnan(info["calibration_scale"]))
        assert np.array_equal(f.signal()[1], reads[1]["signal"])


@pytest.mark.parametrize("pad", [1, 4, 7])
def test_footer_padding_tolerated(tmp_path, pad):
    reads = _reads([1000])
    p = tmp_path / "p.pod5"
    write_pod5(p, reads, footer_pad=pad)
    with D.Pod5File(p) as f:
        assert np.array_equal(f.signal()[0], reads[0]["signal"])


def test_refusals(tmp_path):
    reads = _reads([5000, 300])
    good = tmp_path / "good.pod5"
    write_pod5(good, reads)
    raw = open(good, "rb").read()

    def bad(data, match, name="bad.pod5"):
        q = tmp_path / name
        q.write_bytes(data)
        with pytest.raises(D.DtwaxError, match=match):
            D.Pod5File(q)

    bad(b"\x00" + raw[1:], "signature")
    bad(raw[:-40], "signature")
    bad(raw[:-24] + bytes(16) + raw[-8:], "section markers")
    bad(raw[:8] + raw[24:40] + raw[24:], "section markers|signature|embedded|Arrow")  # shifted content
    with pytest.raises(D.DtwaxError, match="cannot open"):
        D.Pod5File(tmp_path / "missing.pod5")

    def written(match, **kw):
        q = tmp_path / "w.pod5"
        write_pod5(q, reads, **kw)
        with pytest.raises(D.DtwaxError, match=match):
            D.Pod5File(q)

    written("names signal row", bad_row=(1, 0))
    written("num_samples", num_samples_delta=(0, 1))
    written("differs from its 'samples'", vbz=False, samples_delta=(0, 1))
    written("file_identifier", table_file_id=str(uuid.UUID(int=5)))
    written("lacks 'read_id' or 'signal'", omit=("read_id",))
    written("other than ZSTD", ipc_compression="lz4")
    # a VBZ row whose count disagrees with its block fails when decoded
    q = tmp_path / "v.pod5"
    write_pod5(q, reads, samples_delta=(0, -1), num_samples_delta=(0, -1))
    with D.Pod5File(q) as f:
        assert f.signal(reads=[1])[0].size == 300
        with pytest.raises(D.DtwaxError, match="does not hold|zstd"):
            f.signal(reads=[0])


_OPEN_AND_READ = r"""
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import dtwax_amd as D
try:
    with D.Pod5File({path!r}) as f:
        f.signal()
    print("read")
except D.DtwaxError as e:
    print("error:", e)
"""


def test_oversized_lengths_refused(tmp_path):
    """Lengths read from the file cannot overflow the size checks: a plain
    (non-VBZ) file whose signal child node claims 2^62 values and whose row
    offsets claim 2^40 is refused at open (before, it opened and the copy of
    the row crashed the process).  Run in a child so a crash fails the test."""
    import struct
    p = tmp_path / "ok.pod5"
    write_pod5(p, [make_read(np.arange(1000, dtype=np.int16))], vbz=False)
    raw = open(p, "rb").read()
    node = struct.pack("<qq", 1000, 0)  # the signal list's child node: 1000 values, no nulls
    # the row's large_list offsets [0, 1000], followed by the samples 0, 1, 2, 3
    offs = struct.pack("<qq", 0, 1000) + struct.pack("<4h", 0, 1, 2, 3)
    assert raw.count(node) == 1 and raw.count(offs) == 1
    big_node = struct.pack("<qq", 1 << 62, 0)
    big_offs = struct.pack("<qq", 0, 1 << 40) + struct.pack("<4h", 0, 1, 2, 3)
    cases = {"node": raw.replace(node, big_node), "offsets": raw.replace(offs, big_offs),
             "both": raw.replace(node, big_node).replace(offs, big_offs)}
    for name, data in cases.items():
        q = tmp_path / f"{name}.pod5"
        q.write_bytes(data)
        code = _OPEN_AND_READ.format(root=ROOT, tests=os.path.dirname(__file__), path=str(q))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.returncode, r.stderr[-2000:])
        assert r.stdout.startswith("error:"), (name, r.stdout)


def test_signal_window_overflow_refused(tmp_path):
    """dtwax_pod5_signal refuses windows whose sizes overflow (a huge take
    would wrap the destination offsets below the capacity check)."""
    p = tmp_path / "w.pod5"
    write_pod5(p, _reads([3000, 2000]))
    with D.Pod5File(p) as f:
        idx = np.array([0, 1], np.int64)
        n_out = np.zeros(2, np.int64)
        dst = np.zeros(16, np.int16)
        big = (1 << 62) + 8
        rc = D.lib().dtwax_pod5_signal(f._h, idx.ctypes.data, 2, 0, big, dst.ctypes.data, 16,
                                       n_out.ctypes.data, 1)
        assert rc == -1 and b"too small" in D.lib().dtwax_last_error()
        rc = D.lib().dtwax_pod5_signal(f._h, idx.ctypes.data, 2, (1 << 63) - 100, 200, dst.ctypes.data, 16,
                                       n_out.ctypes.data, 1)
        assert rc == -1 and b"overflow" in D.lib().dtwax_last_error()


_FUZZ = r"""
import sys, numpy as np
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import dtwax_amd as D
raw = open({path!r}, 'rb').read()
rng = np.random.default_rng({seed})
ok = err = 0
for it in range({n}):
    b = bytearray(raw)
    for _ in range(int(rng.integers(1, 6))):
        i = int(rng.integers(0, len(b)))
        b[i] = int(rng.integers(0, 256))
    if it % 7 == 0:
        b = b[: int(rng.integers(0, len(b)))]
    open({out!r}, 'wb').write(bytes(b))
    try:
        with D.Pod5File({out!r}) as f:
            f.signal()
            f.signal(skip=100, take=300)
        ok += 1
    except D.DtwaxError:
        err += 1
print(ok, err)
"""


def test_mutated_files_fail_cleanly(tmp_path):
    """Random byte changes and truncations either read or raise DtwaxError;
    the process never crashes (run in a child so a crash fails the test)."""
    reads = _reads([3000, 0, 9000, 50], seed=9)
    p = tmp_path / "f.pod5"
    write_pod5(p, reads, row_samples=2048, signal_batch=2, read_batch=2, shuffle_rows=1)
    code = _FUZZ.format(root=ROOT, tests=os.path.dirname(__file__), path=str(p), seed=1, n=600,
                        out=str(tmp_path / "m.pod5"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    ok, err = map(int, r.stdout.split())
    assert ok + err == 600 and err > 100


def _dump(folder):
    p = subprocess.run([CLI, str(folder) + "/", os.path.join(GOLDEN, "dna_kmer_model.txt"),
                        os.path.join(GOLDEN, "reference.fasta"), "--dump-reads"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout.splitlines(), p.stderr


@pytest.mark.parametrize("vbz", [True, False])
def test_cli_pod5_equals_fast5(tmp_path, vbz):
    """The test read (and a short one) as POD5: the CLI loads the same
    samples under the same name as from the reference's FAST5."""
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    name = "0083a4a2-b062-44ed-988e-e7f5a301121f"
    d = tmp_path / "pod5"
    d.mkdir()
    write_pod5(d / "batch0.pod5", [make_read(sig, 0, read_id=name), make_read(sig[:3000], 1)], vbz=vbz,
               row_samples=1500)
    lines, _ = _dump(d)
    fast5, _ = _dump(GOLDEN)
    rows = [l for l in lines if l.startswith("read_")]
    assert rows == [l for l in fast5 if l.startswith("read_")]
    assert "cuDTW:: Short/invalid reads exempted:: 1" in lines
    assert "cuDTW:: Valid reads (NUM_READS) loaded :: 1" in lines


def test_cli_reports_bad_pod5_and_continues(tmp_path):
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    d = tmp_path / "mix"
    d.mkdir()
    write_pod5(d / "good.pod5", [make_read(sig, 0)])
    (d / "broken.pod5").write_bytes(b"\x8bPOD\r\n\x1a\n" + bytes(100))
    lines, err = _dump(d)
    assert "cuDTW:: Valid reads (NUM_READS) loaded :: 1" in lines
    assert "broken.pod5" in err


@pytest.mark.gpu
def test_cli_scores_pod5_like_fast5(tmp_path):
    """Full CLI run (binary16 sDTW on the GPU) on the POD5 form of the test
    read prints the same table row as on the FAST5."""
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    d = tmp_path / "pod5"
    d.mkdir()
    write_pod5(d / "r.pod5", [make_read(sig, 0, read_id="0083a4a2-b062-44ed-988e-e7f5a301121f")])
    args = [os.path.join(GOLDEN, "dna_kmer_model.txt"), os.path.join(GOLDEN, "reference.fasta")]

    def table(folder):
        p = subprocess.run([CLI, str(folder) + "/", *args], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        return [l for l in p.stdout.splitlines() if "read_" in l and "\t" in l]

    a, b = table(d), table(GOLDEN)
    assert a and a == b


def test_cli_mixed_fast5_and_pod5_folder(tmp_path):
    """A folder with a FAST5 and a POD5 file: the reads of both are loaded
    (each file's reads in its own order), short ones exempted in both."""
    import shutil
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    d = tmp_path / "mixed"
    d.mkdir()
    shutil.copy(os.path.join(GOLDEN, "batch0.fast5"), d / "batch0.fast5")
    write_pod5(d / "more.pod5", [make_read(sig[::-1].copy(), 1), make_read(sig[:1000], 2)])
    lines, _ = _dump(d)
    assert "cuDTW:: 2 ONT fast5 files read" in lines
    assert "cuDTW:: Short/invalid reads exempted:: 1" in lines
    assert "cuDTW:: Valid reads (NUM_READS) loaded :: 2" in lines
    rows = {l.split()[0]: np.array(l.split()[1:], np.float64) for l in lines if l.startswith("read_")}
    assert np.array_equal(rows["read_0083a4a2-b062-44ed-988e-e7f5a301121f"], sig[1952:4000].astype(np.float64))
    rid = "read_" + make_read(sig, 1)["read_id"]
    assert np.array_equal(rows[rid], sig[::-1][1952:4000].astype(np.float64))


@pytest.mark.gpu
def test_classify_pod5_matches_cli(tmp_path):
    """classify_pod5 (Python) and dtwax_main --native (C++) score a POD5 the
    same: same reads kept, same binary16 scores for both strands."""
    import re
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    rng = np.random.default_rng(4)
    other = np.clip(rng.normal(480, 40, 5200), -32768, 32767).astype(np.int16)
    d = tmp_path / "p"
    d.mkdir()
    reads = [make_read(sig, 0, read_id="0083a4a2-b062-44ed-988e-e7f5a301121f"), make_read(sig[:3000], 1),
             make_read(other, 2)]
    write_pod5(d / "r.pod5", reads)
    model, fasta = os.path.join(GOLDEN, "dna_kmer_model.txt"), os.path.join(GOLDEN, "reference.fasta")
    p = subprocess.run([CLI, str(d) + "/", model, fasta, "--native"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    rows = [l.split("\t") for l in p.stdout.splitlines() if re.match(r"^\d+\tread_", l)]
    ctx = D.Context(0)
    try:
        means = D.read_kmer_model(model)
        levels, _ = D.squiggle_build(D.read_fasta(fasta), means, k=6)
        ctx.upload_reference(levels)
        res = D.classify_pod5(ctx, d / "r.pod5", threshold=150.0)
    finally:
        ctx.close()
    assert [r[1] for r in rows] == ["read_" + x for x in res["read_id"]]
    assert list(res["read_id"]) == [reads[0]["read_id"], reads[2]["read_id"]]  # the 3000-sample read is short
    for r, got in zip(rows, res):  # the table prints 6 significant digits, as the reference's log
        assert r[4] == f"{float(got['score'][0]):g}" and r[5] == f"{float(got['score'][1]):g}"
    assert list(res["target"]) == list(res["score"].min(axis=1) < 150.0)
