"""CPU checks of the C ABI library: it builds for gfx950, loads, exports every
symbol include/dtwax.h declares, and its host-only entry points behave."""
import ctypes as C
import os
import re

import numpy as np

import dtwax_amd as D
import oracle as O
from helpers import fasta_python, load_model_means, rev_comp_python

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dtwax.h")).read()
    return sorted(set(re.findall(r"\b(dtwax_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    L = D.lib()
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(D.exported_symbols()) == declared


def test_library_is_gfx950():
    data = open(D.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_squiggle_build_matches_oracle():
    _, means = load_model_means()
    fa = fasta_python()
    lv, skipped = D.squiggle_build(fa, means, ref_len=48128)
    want, sk2 = O.squiggle_reference(fa, rev_comp_python(fa), means, 48128)
    assert skipped == sk2 == 0
    assert np.array_equal(lv, want)


def test_squiggle_build_non_acgt():
    _, means = load_model_means()
    seq = "ACGTACGTNNACGTACGTACnnACGT" * 7
    lv, skipped = D.squiggle_build(seq, means)
    rc = seq[::-1].translate(str.maketrans("ACGT", "TGCA"))  # load_reference.hpp:111-128
    want, sk2 = O.squiggle_reference(seq, rc, means, len(seq) - 5)
    assert skipped == sk2 and np.array_equal(lv, want)


def test_error_paths_without_gpu():
    L = D.lib()
    assert L.dtwax_create(0, None) != 0
    assert L.dtwax_squiggle_build(None, 0, 6, None, 1, None, None) != 0
    p = C.c_void_p()
    rc = L.dtwax_plan_create(None, None, 0, None, C.byref(p))
    assert rc != 0 and L.dtwax_last_error()
