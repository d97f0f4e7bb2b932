"""Threshold search (calc_classification_threshold.py:5-57) on the
reference's own FP16 score tables (src/mapped.log: 397 target reads,
src/unmapped.log: 5,705 non-target reads; values in
tests/golden/fp16_score_logs.npz): 97.62% accuracy for thresholds in
(194.0, 194.125] (the scores are binary16 values 1/8 apart there), and the
reference's grid search returns the first grid value above 194.0."""
import os
import subprocess
import sys

import numpy as np

from helpers import GOLDEN
from dtwax_amd import threshold as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_table(path, rows):
    with open(path, "w") as f:
        f.write("TIMING: 1.5 ms (load_target)\n")
        f.write("Read_ID\tQUERY_LEN\tREF_LEN\tsDTW score: fwd-strand\tsDTW score: rev-strand\n")
        for j, (a, b) in enumerate(rows):
            f.write(f"{j}\tread_{j:08x}\t2048\t48128\t{a:g}\t{b:g}\n")
        f.write("TIMING: 2.0 ms (free)\n")


def test_reference_logs_threshold(tmp_path):
    d = np.load(os.path.join(GOLDEN, "fp16_score_logs.npz"))
    tgt, non = tmp_path / "mapped.log", tmp_path / "unmapped.log"
    _write_table(tgt, d["mapped"])
    _write_table(non, d["unmapped"])
    ts, tl = T.parse_table(tgt)
    ns, nl = T.parse_table(non)
    assert len(ts) == 397 and len(ns) == 5705 and set(tl) == {2048}
    assert np.array_equal(ts, d["mapped"].min(axis=1))
    thr, acc = T.best_threshold(ts, ns)
    # grid step = max score 1545 / 99,999; first grid value above 194.0
    step = 1545.0 / 99999
    assert 194.0 < thr <= 194.0 + step and abs(acc - 97.6237) < 1e-4, (thr, acc)
    # every read has the same length: the per-row form is the same cut / 2048
    thr_r, acc_r = T.best_threshold_per_row(ts, tl, ns, nl)
    assert acc_r == acc and abs(thr_r * 2048 - thr) < 0.02
    out = subprocess.run([sys.executable, "-m", "dtwax_amd.threshold", str(tgt), str(non)],
                         capture_output=True, text=True, cwd=ROOT, check=True).stdout.splitlines()
    assert out[0] == "397" and out[1] == "5705"
    assert out[2].startswith("Best classification threshold: 194.0") and "accuracy: 97.62" in out[2]


def test_threshold_rule_at_equality():
    """target iff score < threshold: a non-target equal to the cut counts as
    correct, a target equal to it does not."""
    thr, acc = T.best_threshold([1.0, 2.0], [3.0, 4.0], num=5)  # grid 0, 1, 2, 3, 4
    assert thr == 3.0 and acc == 100.0
