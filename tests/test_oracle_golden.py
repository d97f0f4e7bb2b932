"""Pins the FP64 oracle and the python-form normalisation against scripts/log.

scripts/log is the reference's saved FP64 trace of scripts/sdtw.py:205-225 for
the test read in test/batch0.fast5 against the lambda reference (fwd[:48128] |
rev[:48128], R = 96,256): the normalised query (4096 values) and D at columns
0, 1, 31, 32, R-2, R-1 for every row, and the final min.  Fixtures were
extracted by tests/golden/make_fixtures.py.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, golden_reference_f64


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "sdtw_log_golden.npz"))


@pytest.fixture(scope="module")
def query(golden):
    sig = np.load(os.path.join(GOLDEN, "batch0_signal.npy"))
    assert sig.dtype == np.int16 and sig.size == 5574
    # scripts/sdtw.py:143-152: signal[1000:1000+4096], discrete_normalize
    return O.python_znorm(sig[1000:1000 + 4096].astype(np.float64))


@pytest.fixture(scope="module")
def reference():
    return golden_reference_f64()


def test_query_normalisation_bit_exact(golden, query):
    assert np.array_equal(query, golden["query"])


def test_every_logged_cell_bit_exact(golden, query, reference):
    assert reference.size == 96256
    tr = O.trace_f64(query, reference, golden["col_index"])
    assert np.array_equal(tr, golden["cols"])


def test_final_min(golden, query, reference):
    sc, am = O.sdtw_f64(query, reference)
    # log prints numpy's 8-significant-digit repr
    assert float(f"{sc:.7f}") == float(golden["final_min"])
    assert abs(sc - 286.8895631) < 5e-8
    assert am == 33305


def test_fp16_logs_are_half_values():
    """src/mapped.log / unmapped.log scores are binary16 values printed with 6
    significant digits: every one must round-trip through float16 within the
    print precision (distribution-level sanity of the FP16 golden)."""
    d = np.load(os.path.join(GOLDEN, "fp16_score_logs.npz"))
    for k in ("mapped", "unmapped"):
        v = d[k].ravel()
        h = v.astype(np.float16).astype(np.float64)
        assert np.all(np.abs(h - v) <= 1e-5 * np.maximum(1.0, np.abs(v)) + 0.5 * np.spacing(np.float16(v)).astype(np.float64))
    assert d["mapped"].shape == (397, 2) and d["unmapped"].shape == (5705, 2)
