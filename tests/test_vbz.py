"""VBZ signal decoding (dtwax_vbz_decode) on the CPU.

Parity unpinned against the VBZ plugin itself: neither the reference nor
this image holds a VBZ-compressed file or the plugin (the reference fetches
vbz_compression v1.0.1 from the network, install_hdf5_vbz.sh:43-58).  What
is pinned: blocks encoded by hand from the published description
(stream-VByte keys and little-endian data, zig-zag differences), zstd by the
system library, round trips against an independent numpy encoder, and
refusal of blocks that do not fit.
"""
import numpy as np
import pytest

import dtwax_amd as D
from vbz_helpers import svb_block, vbz_chunk, zigzag_deltas, zstd_compress


def _hdr(n_samples):
    return int(2 * n_samples).to_bytes(4, "little")


def test_hand_encoded_blocks():
    # samples 5 3 3 300 -2: differences 5 -2 0 297 -302, zig-zag 10 3 0 594 603
    assert list(zigzag_deltas([5, 3, 3, 300, -2])) == [10, 3, 0, 594, 603]
    # version 0 keys (2 bits, 1-4 bytes): 0 0 0 1 | 1 -> 0x40 0x01
    v0 = bytes([0x40, 0x01, 0x0A, 0x03, 0x00, 0x52, 0x02, 0x5B, 0x02])
    assert svb_block([10, 3, 0, 594, 603], "1234") == v0
    got = D.vbz_decode(_hdr(5) + v0, 5, (0, 2, 1, 0))
    assert list(got) == [5, 3, 3, 300, -2]
    # version 1, 0/1/2/4-byte keys: 0 0 0 0 7 -> zig-zag 0 0 0 0 14, keys 0 0 0 0 | 1
    v1 = bytes([0x00, 0x01, 0x0E])
    assert svb_block([0, 0, 0, 0, 14], "0124") == v1
    assert list(D.vbz_decode(_hdr(5) + v1, 5, (1, 2, 1, 0))) == [0, 0, 0, 0, 7]
    # version 1, 1-bit keys (1/2 bytes): 1 200 -> zig-zag 2 398, keys 0b10
    v16 = bytes([0x02, 0x02, 0x8E, 0x01])
    assert svb_block([2, 398], "12") == v16
    assert list(D.vbz_decode(_hdr(2) + v16, 2, (1, 2, 1, 0))) == [1, 200]
    # the same blocks through zstd
    assert list(D.vbz_decode(_hdr(5) + zstd_compress(v0), 5, (0, 2, 1, 1))) == [5, 3, 3, 300, -2]


@pytest.mark.parametrize("version,layout", [(0, "1234"), (1, "0124"), (1, "12")])
@pytest.mark.parametrize("level", [0, 1, 5])
def test_round_trip(version, layout, level):
    rng = np.random.default_rng(version * 10 + level)
    # a nanopore-like random walk with jumps, plus the int16 extremes
    x = np.cumsum(rng.integers(-40, 41, 5000)) + 500
    x[1000] = 32767
    x[1001] = -32768
    x[2000:2010] = x[1999]  # flat run: zero differences
    x = np.clip(x, -32768, 32767).astype(np.int16)
    if layout == "12":  # 1-bit keys hold at most 2 bytes: no full-range jumps
        x[1000:1002] = x[999]
    blob = vbz_chunk(x, version, layout, level)
    got = D.vbz_decode(blob, x.size, (version, 2, 1, level))
    assert np.array_equal(got, x)


def test_refusals():
    x = (np.arange(300) % 50).astype(np.int16)
    good = vbz_chunk(x, 0, "1234", 1)
    for bad, why in [
        (good[:-3], "zstd"),                                     # truncated frame
        (_hdr(301) + good[4:], "fit|zstd"),                      # size header off by one sample
        (good[:2], "shorter"),                                   # no header
        (_hdr(300) + b"\x00" * 40, "zstd frame"),               # not zstd
    ]:
        with pytest.raises(D.DtwaxError, match=why):
            D.vbz_decode(bad, 400, (0, 2, 1, 1))
    with pytest.raises(D.DtwaxError, match="version"):
        D.vbz_decode(good, 400, (2, 2, 1, 1))
    with pytest.raises(D.DtwaxError, match="int16"):
        D.vbz_decode(good, 400, (0, 4, 1, 1))
    with pytest.raises(D.DtwaxError, match="too small"):
        D.vbz_decode(good, 299, (0, 2, 1, 1))
    # a version 0 block read as version 1 fits neither version 1 layout here
    with pytest.raises(D.DtwaxError):
        D.vbz_decode(vbz_chunk(x, 0, "1234", 0), 300, (1, 2, 1, 0))
    # without delta coding the values are the samples themselves
    raw = _hdr(3) + svb_block([7, 0xFFFF, 300], "1234")
    assert list(D.vbz_decode(raw, 3, (0, 2, 0, 0))) == [7, -1, 300]


@pytest.mark.parametrize("vbz", [dict(version=0, layout="1234", zstd_level=1),
                                 dict(version=1, layout="0124", zstd_level=1),
                                 dict(version=1, layout="12", zstd_level=1),
                                 None])
def test_fast5_reader_without_plugin(tmp_path, vbz):
    """The CLI's FAST5 reader on a multi-read file whose Raw/Signal is stored
    in VBZ chunks (no HDF5 plugin in this image): the selected samples equal
    the signal written, chunk boundaries and the zero-filled last chunk
    included."""
    import os
    import subprocess
    from helpers import GOLDEN
    from vbz_helpers import write_fast5
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = os.path.join(root, "dtwax_amd", "bin", "dtwax_main")
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    name = "read_0083a4a2-b062-44ed-988e-e7f5a301121f"
    attrs = dict(digitisation=8192.0, offset=6.0, range=1467.6, sampling_rate=4000.0)
    write_fast5(tmp_path / "vbz.fast5", [(name, sig, attrs), ("read_second", sig[::-1].copy(), attrs)],
                vbz=vbz, chunk=1500)
    p = subprocess.run([cli, str(tmp_path) + "/", os.path.join(GOLDEN, "dna_kmer_model.txt"),
                        os.path.join(GOLDEN, "reference.fasta"), "--dump-reads"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rows = {l.split()[0]: np.array(l.split()[1:], np.float64)
            for l in p.stdout.splitlines() if l.startswith("read_")}
    assert np.array_equal(rows[name], sig[1952:4000].astype(np.float64))
    assert np.array_equal(rows["read_second"], sig[::-1][1952:4000].astype(np.float64))
