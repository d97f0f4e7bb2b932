"""The CPU-runnable FP32 configuration: a 10 kb reference (2 x 9,995
levels), 64 reads x 2000 samples (32 target, 32 from unrelated sequence),
median/MAD normalised; output per read: min cost over the last DP row and
its position.

CPU test: the oracle's scores classify the reads, and a target read's
best column on its source strand lies where the read's k-mers end: its
first k-mer index plus the k-mers it spans (the dwell is geometric with
mean 9 samples, so 2000 samples cover about 222 k-mers), give or take the
dwell spread.  GPU test: the HIP binary32 path equals the oracle bit for
bit, scores and positions.
"""
import numpy as np
import pytest

import oracle as O
from dtwax_amd import synth

SEED, REF_BP, N_READS, QLEN = 1, 10_000, 64, 2000


@pytest.fixture(scope="module")
def c1():
    w = synth.config_workload(seed=SEED, ref_bp=REF_BP, n_reads=N_READS, raw_len=QLEN,
                              target_frac=0.5)
    qs = [synth.medmad_z(r) for r in w.reads]
    ref = w.ref_levels.astype(np.float32)
    sc, am = O.batch_f32(qs, ref)
    return w, qs, ref, sc, am


def test_c1_oracle_classes_and_positions(c1):
    w, qs, ref, sc, am = c1
    assert ref.shape == (2, REF_BP - 5)
    best = sc.min(axis=1)
    # median/MAD of a 2000-sample read misjudges the scale of a few targets,
    # so the classes overlap a little: best single threshold >= 90% accurate
    cuts = np.sort(best)
    acc = max(((best[w.is_target] < t).mean() + (best[~w.is_target] >= t).mean()) / 2 for t in cuts)
    assert acc >= 0.9
    for i in np.flatnonzero(w.is_target):
        strand, start = w.origin[i]
        assert sc[i, strand] <= sc[i, 1 - strand]
        # ~QLEN/9 k-mers covered; allow 4 sigma of the summed geometric dwell
        span = QLEN / 9.0
        sd = np.sqrt(QLEN / 9.0 * (1 - 1 / 9.0) * 81.0) / 9.0
        assert abs(am[i, strand] - (start + span)) < 4 * sd + 10, (i, am[i], start)


@pytest.mark.gpu
def test_c1_gpu_fp32_bit_exact(c1):
    import dtwax_amd as D
    w, qs, ref, sc, am = c1
    ctx = D.Context(0)
    try:
        ctx.upload_reference(ref)
        gsc, gam = ctx.score(qs, precision="f32")
    finally:
        ctx.close()
    assert np.array_equal(gsc, sc)
    assert np.array_equal(gam, am)
