"""Classification threshold from score tables.

Same rule and search as the reference's src/ReadUntil/calc_classification_threshold.py
(:5-34 parse, :49-57 search): a row's score is min(fwd, rev) (fields 4 and 5),
a read is a target iff its score < threshold, and the threshold is the first
of 100,000 evenly spaced values in [0, max score] with the highest accuracy
(correct target + correct non-target calls over all reads).  The sweep is
vectorised (sorted scores and binary search instead of a Python loop).

Besides the reference's fixed-length form, `best_threshold_per_row` fits a
length-aware threshold (score per query row) for reads of different lengths,
as used by the Read Until engine's ``threshold_per_row``.

Command line (the reference's arguments and output lines):

    python -m dtwax_amd.threshold <target_table> <nontarget_table> [--per-row]
"""
from __future__ import annotations

import argparse

import numpy as np


def parse_table(path):
    """(scores, query_lengths) of every table row in a dtwax_main (or
    reference) log: lines with at least 6 fields whose 5th and 6th parse as
    numbers; the header (first field "Read_ID"), TIMING and other log lines
    are skipped."""
    scores, lens = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) < 6 or p[0] == "Read_ID":
                continue
            try:
                fwd, rev, q = float(p[4]), float(p[5]), int(p[2])
            except ValueError:
                continue
            scores.append(min(fwd, rev))
            lens.append(q)
    return np.array(scores, np.float64), np.array(lens, np.int64)


def best_threshold(target, nontarget, num=100_000):
    """Returns (threshold, accuracy in percent) over np.linspace(0, max, num)."""
    t = np.sort(np.asarray(target, np.float64))
    n = np.sort(np.asarray(nontarget, np.float64))
    grid = np.linspace(0, max(t.max(), n.max()), num=num)
    correct = np.searchsorted(t, grid, side="left") + (n.size - np.searchsorted(n, grid, side="left"))
    acc = correct * 100.0 / (t.size + n.size)
    i = int(np.argmax(acc))
    return float(grid[i]), float(acc[i])


def best_threshold_per_row(target, target_lens, nontarget, nontarget_lens, num=100_000):
    """The same search on scores divided by query length (rows)."""
    return best_threshold(np.asarray(target) / np.asarray(target_lens),
                          np.asarray(nontarget) / np.asarray(nontarget_lens), num)


def main(argv=None):
    ap = argparse.ArgumentParser(description="classification threshold from two score tables")
    ap.add_argument("target_file", help="table of target (microbial) reads")
    ap.add_argument("nontarget_file", help="table of non-target (human) reads")
    ap.add_argument("--per-row", action="store_true", help="threshold on score / query length")
    a = ap.parse_args(argv)
    ts, tl = parse_table(a.target_file)
    ns, nl = parse_table(a.nontarget_file)
    print(len(ts))
    print(len(ns))
    if a.per_row:
        thr, acc = best_threshold_per_row(ts, tl, ns, nl)
    else:
        thr, acc = best_threshold(ts, ns)
    print(f"Best classification threshold: {thr}, accuracy: {acc}")


if __name__ == "__main__":
    main()
