"""Pins the oracle's statistics: numpy summation order, normalisers, and the
reference squiggle build (load_reference.hpp:33-108 arithmetic)."""
import numpy as np

import oracle as O
from helpers import fasta_python, load_model_means, rev_comp_python


def test_numpy_sum_order():
    rng = np.random.default_rng(21)
    for n in list(range(0, 200)) + [255, 256, 1000, 4096, 8191, 8192, 8193, 16385, 48128, 100001]:
        a = rng.standard_normal(n) * 1e3
        assert O.numpy_sum(a) == np.sum(a)
        if n:
            assert O.numpy_sum(a) / n == np.mean(a)


def test_python_znorm_matches_numpy():
    rng = np.random.default_rng(22)
    x = rng.integers(300, 900, 4096).astype(np.float64)
    mean = np.mean(x)
    # libm pow(v, 0.5): the form scripts/log pins bit-exactly.  (numpy 2.x
    # evaluates np.power on a float64 scalar with its own SIMD pow, which can
    # differ from libm in the last bit.)
    import math
    sd = math.pow(float(np.mean(np.power(x - mean, 2))), 0.5)
    assert np.array_equal(O.python_znorm(x), np.array([(i - mean) / sd for i in x]))


def test_meanstd_closed_forms():
    rng = np.random.default_rng(23)
    x = (rng.standard_normal(2048) * 17 + 500).astype(np.float32)
    y, m, s = O.norm_meanstd(x, eps=0.0)
    assert abs(np.mean(y.astype(np.float64))) < 1e-6
    assert abs(np.mean(y.astype(np.float64) ** 2) - 1.0) < 1e-5
    y2, _, s2 = O.norm_meanstd(x, eps=1e-4)
    assert s2 > s and np.all(np.abs(y2) <= np.abs(y) + 1e-6)
    # constant input: eps keeps it finite, all zeros
    y3, _, _ = O.norm_meanstd(np.full(64, 7.0, np.float32), eps=1e-4)
    assert np.all(y3 == 0)


def test_medmad_matches_numpy():
    rng = np.random.default_rng(24)
    for n in (1, 2, 3, 10, 1999, 2000, 4096):
        x = (rng.standard_normal(n) * 9 + 80).astype(np.float32)
        y, med, mad = O.norm_medmad(x)
        xm = np.median(x.astype(np.float64))
        assert med == xm
        assert mad == np.median(np.abs(x.astype(np.float64) - xm))
        if mad > 0:
            want = ((x.astype(np.float64) - xm) / (1.4826 * mad)).astype(np.float32)
            assert np.allclose(y, want, rtol=0, atol=2e-7 * max(1.0, np.max(np.abs(want))))


def _squiggle_numpy_f32(seq, model, ref_len, k=6):
    """Independent restatement with numpy float32 scalars, sequential sums."""
    lv = []
    for i in range(ref_len):
        km = seq[i:i + k]
        lv.append(np.float32(model[km]) if km in model else None)
    mean, sq = np.float32(0), np.float32(0)
    for v in lv:
        if v is not None:
            mean = np.float32(mean + v)
            sq = np.float32(sq + np.float32(v * v))
    mean = np.float32(mean / np.float32(ref_len))
    sq = np.float32(sq / np.float32(ref_len))
    sd = np.float32(np.sqrt(np.float32(sq - np.float32(mean * mean))))
    out = np.zeros(ref_len, np.float32)
    for i, v in enumerate(lv):
        if v is not None:
            out[i] = np.float32(np.float32(v - mean) / sd)
    return out


def test_squiggle_reference_lambda():
    model, means = load_model_means()
    fa = fasta_python()
    rc = rev_comp_python(fa)
    ref_len = 3000  # the float arithmetic per element is what is pinned
    out, skipped = O.squiggle_reference(fa, rc, means, ref_len)
    assert skipped == 0
    assert np.array_equal(out[0], _squiggle_numpy_f32(fa, model, ref_len))
    assert np.array_equal(out[1], _squiggle_numpy_f32(rc, model, ref_len))


def test_squiggle_non_acgt_kmers():
    model, means = load_model_means()
    seq = "ACGTACGTNNACGTACGTAC" * 5
    rc = rev_comp_python(seq)  # N stays N
    out, skipped = O.squiggle_reference(seq, rc, means, 60)
    ref = _squiggle_numpy_f32(seq, model, 60)
    assert np.array_equal(out[0], ref)
    assert skipped > 0 and np.all(out[0][[8, 9]] == 0)
