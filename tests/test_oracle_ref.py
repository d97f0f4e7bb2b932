"""The reference's own host code, built from its sources (oracle/Makefile
target `ref`, only where the reference checkout is mounted), against the
product and the oracle:

* load_reference.hpp's squiggle build (binary32 mode: the float levels its
  FP16 build rounds to half) on the lambda fixture equals
  dtwax_squiggle_build and the oracle bit for bit, both strands, all 48,128
  k-mers;
* the reference's FAST5 reader (fast5/nanopolish_fast5_io.cpp) and the CLI's
  reader select the same reads, in the same order, with the same samples
  [1952, 4000), on the fixture and on multi-read files with short reads.
"""
import os
import subprocess

import numpy as np
import pytest

import dtwax_amd as D
import oracle as O
from helpers import GOLDEN, fasta_python, load_model_means, rev_comp_python

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
CLI = os.path.join(ROOT, "dtwax_amd", "bin", "dtwax_main")

need = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_squiggle")),
                          reason="oracle/_ref not built (the reference checkout is not mounted)")


@need
def test_squiggle_equals_reference_loader(tmp_path):
    out = tmp_path / "sq.f32"
    subprocess.run([os.path.join(REF, "ref_squiggle"), os.path.join(GOLDEN, "dna_kmer_model.txt"),
                    os.path.join(GOLDEN, "reference.fasta"), str(out)], check=True, timeout=300,
                   capture_output=True)
    ref = np.fromfile(out, np.float32).reshape(2, 48128)
    seq = fasta_python()
    _, means = load_model_means()
    prod, skipped = D.squiggle_build(seq, means, ref_len=48128)
    assert skipped == 0 and np.array_equal(prod, ref)
    orc, _ = O.squiggle_reference(seq, rev_comp_python(seq), means, 48128)
    assert np.array_equal(orc, ref)


def _cli_reads(folder):
    p = subprocess.run([CLI, str(folder) + "/", os.path.join(GOLDEN, "dna_kmer_model.txt"),
                        os.path.join(GOLDEN, "reference.fasta"), "--dump-reads"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rows = [l.split() for l in p.stdout.splitlines() if l.startswith("read_")]
    short = [l for l in p.stdout.splitlines() if "Short/invalid reads exempted" in l][0]
    return [(r[0], np.array(r[1:], np.float64)) for r in rows], int(short.split("::")[-1])


@need
def test_fast5_reader_equals_reference_reader(tmp_path):
    from vbz_helpers import write_fast5
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.int16)
    rng = np.random.default_rng(3)
    attrs = dict(digitisation=8192.0, offset=6.0, range=1467.6, sampling_rate=4000.0)
    reads = [("read_c", sig, attrs), ("read_a", rng.integers(200, 900, 3000).astype(np.int16), attrs),
             ("read_b", rng.integers(200, 900, 9000).astype(np.int16), attrs),
             ("read_d", rng.integers(200, 900, 4000).astype(np.int16), attrs)]
    write_fast5(tmp_path / "multi.fast5", reads, vbz=None, chunk=2000)
    for folder, files in ((GOLDEN, [os.path.join(GOLDEN, "batch0.fast5")]),
                          (tmp_path, [str(tmp_path / "multi.fast5")])):
        p = subprocess.run([os.path.join(REF, "ref_fast5"), *files], capture_output=True,
                           text=True, timeout=120, check=True)
        ref_rows = [l.split() for l in p.stdout.splitlines()]
        valid = [(r[0], np.array(r[2:], np.float64)) for r in ref_rows if int(r[1]) >= 4000]
        n_short = sum(int(r[1]) < 4000 for r in ref_rows)
        ours, ours_short = _cli_reads(folder)
        assert [g for g, _ in ours] == [g for g, _ in valid]
        assert ours_short == n_short
        for (_, a), (_, b) in zip(ours, valid):
            assert np.array_equal(a, b)
