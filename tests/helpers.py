"""Shared helpers for the test suite (fixture loading, python-form reference)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_model_means(path=None):
    """6-mer model file -> (dict kmer->mean as Python float, means[4096] float32
    indexed by 2-bit code A=0 C=1 G=2 T=3, first base most significant)."""
    path = path or os.path.join(GOLDEN, "dna_kmer_model.txt")
    d = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) >= 2:
                d[p[0]] = float(p[1])
    codes = {"A": 0, "C": 1, "G": 2, "T": 3}
    arr = np.zeros(4096, np.float32)
    for k, v in d.items():
        c = 0
        for ch in k:
            c = c * 4 + codes[ch]
        arr[c] = np.float32(v)
    return d, arr


def fasta_python(path=None):
    """scripts/sdtw.py:60-63 get_fasta: every line after the first, joined."""
    path = path or os.path.join(GOLDEN, "reference.fasta")
    with open(path) as f:
        return "".join(f.read().split("\n")[1:])


def rev_comp_python(bases):
    """scripts/sdtw.py:66-68."""
    return bases.replace("A", "t").replace("T", "a").replace("G", "c").replace("C", "g").upper()[::-1]


def golden_reference_f64():
    """fwd[:48128] | rev[:48128], each python-z-normalised (scripts/sdtw.py:101-116, 234)."""
    import oracle as O
    model, _ = load_model_means()
    fa = fasta_python()

    def sig(seq):
        v = np.array([model[seq[i:i + 6]] for i in range(len(seq) - 6 + 1)], np.float64)
        return O.python_znorm(v[:48128])

    return np.concatenate([sig(fa), sig(rev_comp_python(fa))])
