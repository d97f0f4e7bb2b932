"""The CLI's host code under AddressSanitizer and UBSan (CPU only).

`make -C dtwax_amd/csrc asan` builds dtwax_main_asan: the drop-in driver's
argument parser, FASTA and k-mer model loaders, FAST5 reader, VBZ decoder
and POD5 reader compiled with -fsanitize=address,undefined (its
--dump-reads mode never touches the GPU).  It reads a committed corpus of
mutated POD5 and VBZ FAST5 files (tests/golden/fuzz_corpus.npz, written by
tests/golden/make_fuzz_corpus.py): every file either reads or is reported
and skipped, with no sanitizer report; the unchanged control files' reads
come through whole.  Malformed command lines exit with status 2, cleanly.
"""
import fcntl
import os
import subprocess
import sys
import tempfile

import pytest

from helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dtwax_amd", "csrc")
ASAN_BIN = os.path.join(ROOT, "dtwax_amd", "bin", "dtwax_main_asan")
MODEL = os.path.join(GOLDEN, "dna_kmer_model.txt")
FASTA = os.path.join(GOLDEN, "reference.fasta")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def asan_bin():
    # parallel test workers each reach this fixture: one builds (or relinks)
    # while the others wait, so none runs a half-written binary
    with open(os.path.join(tempfile.gettempdir(), "dtwax_asan_build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        r = subprocess.run(["make", "-C", CSRC, "asan"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return ASAN_BIN


def _clean(p):
    return "AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr


def test_mutated_corpus_under_sanitizers(asan_bin, tmp_path):
    sys.path.insert(0, GOLDEN)
    from make_fuzz_corpus import corpus_files
    files = corpus_files()
    assert len(files) == 62
    for name, data in files.items():
        (tmp_path / name).write_bytes(data)
    p = subprocess.run([asan_bin, str(tmp_path) + "/", MODEL, FASTA, "--dump-reads"], env=ENV,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0 and _clean(p), p.stderr[-4000:]
    rows = [l.split()[0] for l in p.stdout.splitlines() if l.startswith("read_")]
    assert "cuDTW:: 62 ONT fast5 files read" in p.stdout
    # the controls: the POD5's reads of >= 4000 samples (3 of 6) and the
    # FAST5's three reads (>= 4000 samples each) all load
    for c in ("00000000-0000-0000-0000-000000000000", "00000001-0000-0000-0000-000000000000",
              "00000002-0000-0000-0000-000000000000"):
        assert sum(r == f"read_{c}" for r in rows) >= 1, c
    assert len(rows) >= 6


@pytest.mark.parametrize("args", [["--precision", "f64"], ["--query-len"], ["--norm", "x"], ["--bogus"],
                                  ["--query-len", "-3"]])
def test_command_line_errors_under_sanitizers(asan_bin, args):
    p = subprocess.run([asan_bin, GOLDEN + "/", MODEL, FASTA, *args, "--dump-reads"], env=ENV,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and _clean(p), (p.returncode, p.stderr[-2000:])
    p = subprocess.run([asan_bin, MODEL, FASTA, "pipe", *args], env=ENV, cwd=str(GOLDEN),
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and _clean(p), (p.returncode, p.stderr[-2000:])
