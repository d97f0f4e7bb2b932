"""Parity at the workloads' full sizes.

* 48.5 kb reference x 4096 reads x 2000 samples (3000 raw, 1000 trimmed),
  binary16, scored in one launch: a random sample of 512 reads is compared
  bit for bit (scores and argmins) with the oracle; the oracle finishes the
  sample in seconds.
* 5 Mb reference (2 x 4,999,995 columns), reads of 500..8000 samples, as
  the uneven workload plans them (column chunks of 65536, per-read widths):
  properties that do not depend on size -- a read cut from either strand
  scores exactly 0 on that strand at its last column; chunk lengths
  65536 and 262144 give identical results; the score of a prefix never
  exceeds the score of the whole read (the last-row minimum is
  non-decreasing in the number of rows).
"""
import numpy as np
import pytest

import dtwax_amd as D
import oracle as O
from dtwax_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def test_48kb_4096_reads_sample_bit_exact(ctx):
    w = synth.config_workload(seed=2, ref_bp=48_500, n_reads=4096, raw_len=3000, target_frac=0.5)
    reads = synth.trim_reads(w.reads, 1000)
    norm, _ = ctx.normalize(reads, mode=D.NORM_MEDMAD)
    ctx.upload_reference(w.ref_levels)
    print("48.5 kb x 4096 x 2000, f16, one launch", flush=True)
    sc, am = ctx.score(norm, precision="f16")
    pick = np.sort(np.random.default_rng(0).choice(len(norm), 512, replace=False))
    qs = [O.f32_to_f16(norm[i]) for i in pick]
    sh = np.stack([O.f32_to_f16(w.ref_levels[k]) for k in range(2)])
    osc, oam = O.batch_f16(qs, sh)
    osc = O.f16_to_f32(osc.ravel()).reshape(osc.shape)
    assert np.array_equal(sc[pick], osc) and np.array_equal(am[pick], oam)


def _plan_scores(ctx, reads, **kw):
    flat, offs, lens = D.pack_reads(reads)
    p = ctx.plan(lens, **kw)
    d_q, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lens))
    d_out = D.DeviceArray((len(reads),), D.RESULT_DTYPE)
    p.run(d_q, d_off, d_len, d_out, None)
    res = d_out.to_host()
    p.close()
    return res["score"].copy(), res["argmin"].copy()


def test_5mb_reference_properties(ctx):
    rng = np.random.default_rng(21)
    means, _ = synth.make_model(rng)
    _, levels = synth.build_reference(rng, 5_000_000, means)
    # binary16 values, so a cut of the reference is an exact query
    levels = O.f16_to_f32(O.f32_to_f16(levels.ravel())).reshape(levels.shape)
    ctx.upload_reference(levels)
    N = levels.shape[1]
    lens = rng.integers(500, 8001, 48)
    planted = []
    reads = []
    for i, L in enumerate(lens[:24]):
        strand, start = i % 2, int(rng.integers(0, N - L))
        reads.append(levels[strand, start:start + L].copy())
        planted.append((strand, start + L - 1))
    for L in lens[24:]:
        reads.append(rng.standard_normal(int(L)).astype(np.float32))
    print("5 Mb, 48 reads, chunk 65536", flush=True)
    sc, am = _plan_scores(ctx, reads, precision="f16", segment=0, ref_chunk=65536)
    for i, (strand, end) in enumerate(planted):
        assert sc[i, strand] == 0.0 and am[i, strand] == end, (i, sc[i], am[i])
        assert sc[i, 1 - strand] > 0.0
    print("5 Mb, 48 reads, chunk 262144", flush=True)
    sc2, am2 = _plan_scores(ctx, reads, precision="f16", segment=0, ref_chunk=262144)
    assert np.array_equal(sc, sc2) and np.array_equal(am, am2)
    print("5 Mb, prefixes", flush=True)
    halves = [r[: r.size // 2] for r in reads[24:]]
    sch, _ = _plan_scores(ctx, halves, precision="f16", segment=0, ref_chunk=65536)
    assert np.all(sch <= sc[24:])
