"""The drop-in command line driver: same arguments, read selection and log
format as the reference's offline main (src/main.cu:53-318)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, fasta_python, load_model_means

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dtwax_amd", "bin", "dtwax_main")
ARGS = [GOLDEN + "/", os.path.join(GOLDEN, "dna_kmer_model.txt"), os.path.join(GOLDEN, "reference.fasta")]


def run(*extra, check=True):
    p = subprocess.run([CLI, *ARGS, *extra], capture_output=True, text=True, timeout=300)
    if check:
        assert p.returncode == 0, p.stderr
    return p


def test_dump_reads_selection_and_trim():
    """Valid reads keep samples [1952, 4000) as raw ADC values."""
    p = run("--dump-reads")
    lines = p.stdout.splitlines()
    assert "cuDTW:: Loading ONT fast5 reads from " + GOLDEN + "/" in lines
    assert "cuDTW:: 1 ONT fast5 files read" in lines
    assert "cuDTW:: Short/invalid reads exempted:: 0" in lines
    assert "cuDTW:: Valid reads (NUM_READS) loaded :: 1" in lines
    row = [l for l in lines if l.startswith("read_")][0].split()
    assert row[0] == "read_0083a4a2-b062-44ed-988e-e7f5a301121f"
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy"))
    assert np.array_equal(np.array(row[1:], np.float64), sig[1952:4000].astype(np.float64))


def test_short_reads_are_exempted():
    p = run("--dump-reads", "--query-len", "4000")
    assert "cuDTW:: Short/invalid reads exempted:: 1" in p.stdout.splitlines()
    assert "cuDTW:: Valid reads (NUM_READS) loaded :: 0" in p.stdout.splitlines()


def test_usage_and_errors():
    p = subprocess.run([CLI], capture_output=True, text=True)
    assert p.returncode == 2 and "usage" in p.stderr
    p = subprocess.run([CLI, GOLDEN + "/", ARGS[1], "/nonexistent.fasta", "--dump-reads"],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "Error opening" in p.stderr


def _lambda_squiggle(n=48128):
    _, means = load_model_means()
    fa = fasta_python()
    rc = fa[::-1].translate(str.maketrans("ACGT", "TGCA"))
    lv, _ = O.squiggle_reference(fa, rc, means, n)
    return lv


def _test_read_query():
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.float32)[1952:4000]
    q, _, _ = O.norm_meanstd(sig, 1e-4)
    return q


@pytest.mark.gpu
def test_table_matches_oracle():
    """Full run on the test read: by default the table rows carry the
    reference kernel's own binary16 scores (the literal emulator of DTW.cu,
    seam deviations included) over the 47,424 scored columns of the
    48,128-k-mer squiggle (mean/std normalised query, eps 1e-4); with --clean
    the clean recurrence's."""
    p = run()
    lines = p.stdout.splitlines()
    hdr = "Read_ID\tQUERY_LEN\tREF_LEN\tsDTW score: fwd-strand\tsDTW score: rev-strand"
    assert hdr in lines
    assert any(re.match(r"TIMING: \S+ ms \S+ Mbps \(concurrent_DTW_kernel_launch\)", l) for l in lines)
    rows = [l.split("\t") for l in lines if re.match(r"^\d+\tread_", l)]
    assert len(rows) == 1 and rows[0][1] == "read_0083a4a2-b062-44ed-988e-e7f5a301121f"
    assert rows[0][2] == "2048" and rows[0][3] == "48128"
    fwd, rev = float(rows[0][4]), float(rows[0][5])
    lv = _lambda_squiggle()
    q = _test_read_query()
    flat = O.f16_to_f32(O.f32_to_f16(lv.ravel())).reshape(2, -1)
    want = O.emulate_reference(flat, q[None, :])[0]
    assert float(f"{want[0]:g}") == fwd and float(f"{want[1]:g}") == rev, (fwd, rev, want)
    # --clean: the clean recurrence on the same launch
    p = run("--clean")
    row = [l.split("\t") for l in p.stdout.splitlines() if re.match(r"^\d+\tread_", l)][0]
    qh = O.f32_to_f16(q)
    for k in range(2):
        sc, _ = O.sdtw_f16(qh, O.f32_to_f16(lv[k, :47424]))
        w = float(O.f16_to_f32(np.array([sc], np.uint16))[0])
        assert float(f"{w:g}") == float(row[4 + k]), (k, row, w)


@pytest.mark.gpu
def test_read_until_fifo_matches_oracle(tmp_path):
    """The Read Until form (src/ru_main.cu:172-326): 10 x 2048 bytes per batch
    from the FIFO `pipe` in the working directory, one unsigned sample per
    byte, mean/std normalised, rows `j j 2048 48128 fwd rev` without a
    header; the loop ends when the writer closes the FIFO.  Scores are the
    reference kernel's own (the same kernel as the offline driver)."""
    import errno
    import time
    fifo = tmp_path / "pipe"
    os.mkfifo(fifo)
    p = subprocess.Popen([CLI, ARGS[1], ARGS[2], "pipe"], cwd=tmp_path, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    data = np.random.default_rng(5).integers(0, 256, 2 * 20480, dtype=np.uint8)
    deadline = time.time() + 120
    fd = None
    while fd is None:  # the driver opens the FIFO once its reference is built
        try:
            fd = os.open(fifo, os.O_WRONLY | os.O_NONBLOCK)
        except OSError as e:
            assert e.errno == errno.ENXIO and p.poll() is None and time.time() < deadline, \
                (e, p.poll(), p.stderr.read() if p.poll() is not None else "")
            time.sleep(0.05)
    os.set_blocking(fd, True)
    os.write(fd, data.tobytes())
    os.close(fd)
    out, err = p.communicate(timeout=120)
    assert p.returncode == 0, err
    rows = [l.split("\t") for l in out.splitlines() if l.strip()]
    assert len(rows) == 20 and all(len(r) == 6 for r in rows)
    lv = _lambda_squiggle()
    flat = O.f16_to_f32(O.f32_to_f16(lv.ravel())).reshape(2, -1)
    check = [0, 1, 2, 10, 11, 12]  # rows of both batches, through the emulator
    qs = []
    for i in check:
        b, j = divmod(i, 10)
        sig = data[b * 20480 + j * 2048:b * 20480 + (j + 1) * 2048].astype(np.float32)
        q, _, _ = O.norm_meanstd(sig, 1e-4)
        qs.append(q)
    want = O.emulate_reference(flat, np.stack(qs))
    for i, r in enumerate(rows):
        b, j = divmod(i, 10)
        assert r[:4] == [str(j), str(j), "2048", "48128"]
        if i in check:
            w = want[check.index(i)]
            for k in range(2):
                assert float(f"{w[k]:g}") == float(r[4 + k]), (i, k, r, w)


@pytest.mark.gpu
def test_fp32_build_and_native_modes():
    """--fp32-build: the reference's FP32 build (one concatenated fwd|rev
    sequence of 2 x 48,128 levels, 115 tiles of 832 scored, one score column)
    with its kernel's own DP (the emulator in its FP32 mode), and with
    --clean the clean binary32 recurrence; --native --precision f32 --norm
    medmad: every k-mer, both strands, binary32."""
    _, means = load_model_means()
    fa = fasta_python()
    rc = fa[::-1].translate(str.maketrans("ACGT", "TGCA"))
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy")).astype(np.float32)[1952:4000]
    # FP32 build
    p = run("--fp32-build")
    lines = p.stdout.splitlines()
    assert "Read_ID\tQUERY_LEN\tREF_LEN\tsDTW-score" in lines
    row = [l.split("\t") for l in lines if re.match(r"^\d+\tread_", l)][0]
    assert row[2] == "2048" and row[3] == "96256" and len(row) == 5
    lv, _ = O.squiggle_reference(fa, rc, means, 48128)
    full = np.concatenate([lv[0], lv[1]])
    q, _, _ = O.norm_meanstd(sig, 1e-4)
    want = O.emulate_reference(full, q[None, :], ref_len=2 * 48128, fp16=False)[0]
    assert float(f"{want:g}") == float(row[4]), (row, want)
    p = run("--fp32-build", "--clean")
    row = [l.split("\t") for l in p.stdout.splitlines() if re.match(r"^\d+\tread_", l)][0]
    want, _ = O.sdtw_f32(q, full[: (2 * 48128 // 832) * 832])
    assert float(f"{want:g}") == float(row[4]), (row, want)
    # native, binary32, median/MAD
    p = run("--native", "--precision", "f32", "--norm", "medmad")
    row = [l.split("\t") for l in p.stdout.splitlines() if re.match(r"^\d+\tread_", l)][0]
    n_k = len(fa) - 5
    lvn, _ = O.squiggle_reference(fa, rc, means, n_k)
    qm, _, _ = O.norm_medmad(sig)
    assert row[3] == str(n_k)
    for k in range(2):
        w, _ = O.sdtw_f32(qm, lvn[k])
        assert float(f"{w:g}") == float(row[4 + k]), (k, row, w)
