"""Writes tests/golden/fuzz_corpus.npz: the mutated input files that
tests/test_host_sanitized.py feeds to the CLI's readers built with
AddressSanitizer and UBSan.

Contents (seeded, reproducible with `python tests/golden/make_fuzz_corpus.py`):
* POD5 files of the test writer (tests/pod5_helpers.py), VBZ and plain
  signal rows, small batches, shuffled rows, then 1-6 random byte changes
  anywhere (every fifth file also truncated at a random point);
* the two oversized-length cases of the POD5 reader's regression test (a
  signal child node of 2^62 values, row offsets of 2^40);
* multi-read FAST5 files whose Raw/Signal is stored in VBZ chunks (written
  with tests/vbz_helpers.py, no HDF5 plugin), with 1-6 random byte changes
  inside the chunk payloads only (HDF5's own structure stays intact, so the
  reads reach this build's VBZ decoder);
* one unchanged POD5 and one unchanged FAST5 file as controls.
Stored compactly: the unchanged base files (`base_<i>`, uint8), and per
corpus file its name, base, truncation length (-1: none) and byte changes
(`edits`: rows of file index, position, new value); `corpus_files` rebuilds
the files.
"""
import os
import struct
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from pod5_helpers import make_read, write_pod5  # noqa: E402
from vbz_helpers import vbz_chunk, write_fast5  # noqa: E402

OUT = os.path.join(HERE, "fuzz_corpus.npz")


def _signals(rng, lens):
    return [np.cumsum(rng.integers(-40, 41, n)).astype(np.int16) + 500 for n in lens]


def _mutate(raw, rng, lo=0, hi=None, truncate=False):
    """(byte changes [(pos, value)], truncation length or -1)"""
    hi = len(raw) if hi is None else hi
    edits = [(int(rng.integers(lo, hi)), int(rng.integers(0, 256))) for _ in range(int(rng.integers(1, 7)))]
    return edits, int(rng.integers(0, len(raw))) if truncate else -1


def corpus_files(path=OUT):
    """{file name: bytes} rebuilt from the stored bases and edits."""
    d = np.load(path)
    bases = [d[f"base_{i}"].tobytes() for i in range(int(d["n_bases"]))]
    out = {}
    for fi, (name, base, trunc) in enumerate(zip(d["names"], d["base_of"], d["trunc"])):
        b = bytearray(bases[int(base)])
        for _, pos, val in d["edits"][d["edits"][:, 0] == fi]:
            b[int(pos)] = int(val)
        out[str(name)] = bytes(b if trunc < 0 else b[: int(trunc)])
    return out


def main():
    rng = np.random.default_rng(2024)
    bases, files = [], {}  # files: name -> (base index, edits, truncation)

    def base(raw):
        bases.append(raw)
        return len(bases) - 1
    with tempfile.TemporaryDirectory() as tmp:
        lens = [4200, 0, 9000, 50, 2100, 4001]
        reads = [make_read(s, i) for i, s in enumerate(_signals(rng, lens))]
        for vbz in (True, False):
            p = os.path.join(tmp, "base.pod5")
            write_pod5(p, reads, vbz=vbz, row_samples=2048, signal_batch=3, read_batch=2, shuffle_rows=5)
            raw = open(p, "rb").read()
            bi = base(raw)
            if vbz:
                files["control.pod5"] = (bi, [], -1)
            for k in range(20):
                files[f"pod5_{'vbz' if vbz else 'plain'}_{k:02d}.pod5"] = (bi, *_mutate(raw, rng, truncate=k % 5 == 4))
        # oversized lengths (a plain file of one 1000-sample read)
        p = os.path.join(tmp, "one.pod5")
        write_pod5(p, [make_read(np.arange(1000, dtype=np.int16))], vbz=False)
        raw = open(p, "rb").read()
        node, offs = struct.pack("<qq", 1000, 0), struct.pack("<qq", 0, 1000) + struct.pack("<4h", 0, 1, 2, 3)
        assert raw.count(node) == 1 and raw.count(offs) == 1
        bi = base(raw)
        big_node, big_offs = struct.pack("<qq", 1 << 62, 0), struct.pack("<q", 1 << 40)
        files["oversized_node.pod5"] = (bi, [(raw.find(node) + j, big_node[j]) for j in range(16)], -1)
        files["oversized_offsets.pod5"] = (bi, [(raw.find(offs) + 8 + j, big_offs[j]) for j in range(8)], -1)
        # VBZ FAST5: mutate inside the stored chunks only
        sigs = _signals(rng, [5000, 4100, 6000])
        attrs = dict(digitisation=8192.0, offset=6.0, range=1467.6, sampling_rate=4000.0)
        recs = [(f"read_{i:08d}-0000-0000-0000-000000000000", s, attrs) for i, s in enumerate(sigs)]
        for fi, vbz in enumerate([dict(version=0, zstd_level=1), dict(version=1, layout="0124", zstd_level=1),
                                  dict(version=0, zstd_level=0)]):
            p = os.path.join(tmp, "base.fast5")
            chunk = 1500
            write_fast5(p, recs, vbz=vbz, chunk=chunk)
            raw = open(p, "rb").read()
            bi = base(raw)
            if fi == 0:
                files["control.fast5"] = (bi, [], -1)
            spans = []
            for _, s, _ in recs:
                for off in range(0, s.size, chunk):
                    part = np.zeros(chunk, np.int16)
                    part[: min(chunk, s.size - off)] = s[off:off + chunk]
                    blob = vbz_chunk(part, vbz.get("version", 0), vbz.get("layout"), vbz.get("zstd_level", 1))
                    at = raw.find(blob)
                    assert at >= 0
                    spans.append((at, at + len(blob)))
            for k in range(6):
                a, z = spans[int(rng.integers(0, len(spans)))]
                files[f"fast5_vbz{fi}_{k:02d}.fast5"] = (bi, *_mutate(raw, rng, lo=a, hi=z))
    names = sorted(files)
    edits = [(fi, pos, val) for fi, n in enumerate(names) for pos, val in files[n][1]]
    np.savez_compressed(OUT, n_bases=np.int64(len(bases)), names=np.array(names),
                        base_of=np.array([files[n][0] for n in names], np.int64),
                        trunc=np.array([files[n][2] for n in names], np.int64),
                        edits=np.array(edits, np.int64).reshape(-1, 3),
                        **{f"base_{i}": np.frombuffer(b, np.uint8) for i, b in enumerate(bases)})
    print(f"{OUT}: {len(names)} files from {len(bases)} bases, {len(edits)} byte changes")


if __name__ == "__main__":
    main()
