"""Python FASTA and k-mer model readers (dtwax_amd.read_fasta,
read_kmer_model) against the fixture files and the rules of
load_reference.hpp:131-212."""
import os

import numpy as np
import pytest

import dtwax_amd as D
from helpers import GOLDEN, fasta_python, load_model_means


def test_fixture_files():
    seq = D.read_fasta(os.path.join(GOLDEN, "reference.fasta"))
    assert seq == fasta_python() and len(seq) == 48502
    means = D.read_kmer_model(os.path.join(GOLDEN, "dna_kmer_model.txt"))
    assert means.size == 4096 and np.array_equal(means, load_model_means()[1])


def test_fasta_rules(tmp_path):
    p = tmp_path / "a.fa"
    p.write_text(">one\nACGT\nTTAA\r\n>two\nGG\n")
    assert D.read_fasta(p) == "GG"  # the last entry is used
    p.write_text(">one\nAC GT\nTT\n")  # a sequence line with a space drops the entry
    with pytest.raises(ValueError):
        D.read_fasta(p)
    p.write_text("ACGT\n")
    with pytest.raises(ValueError):
        D.read_fasta(p)


def test_model_rules(tmp_path):
    p = tmp_path / "m.txt"
    p.write_text("kmer\tlevel_mean\tlevel_stdv\nAA 1.5 0.1\nAC 2.5 0.1\nAA 9.0 0.1\nTT 4.0 0.2\n")
    m = D.read_kmer_model(p)
    assert m.size == 16 and m[0] == 1.5 and m[1] == 2.5 and m[15] == 4.0  # first AA wins
    p.write_text("\n")
    with pytest.raises(ValueError):
        D.read_kmer_model(p)
