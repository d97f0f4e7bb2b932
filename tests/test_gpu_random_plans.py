"""Randomised plan layouts against the oracle (seeded, so every run checks
the same cases): read lengths, segment widths, reference chunking, packing
and the two binary16 forms in combination.  Scores and argmins bit-exact."""
import numpy as np
import pytest

import dtwax_amd as D
import oracle as O

pytestmark = pytest.mark.gpu

WIDTHS = [0, 8, 12, 16, 20, 25, 28, 32]


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    n_cols = int(rng.integers(80, 2500))
    n_reads = int(rng.integers(1, 24))
    # a mix of tiny, one-band and multi-band reads
    lens = [int(x) for x in np.where(rng.random(n_reads) < 0.2, rng.integers(1, 70, n_reads),
                                     rng.integers(70, 5000, n_reads))]
    segment = int(rng.choice(WIDTHS))
    precision = "f16_fused" if rng.random() < 0.3 else "f16"
    packing = int(rng.choice([-1, 0, 1]))
    if packing == 1 and segment and segment < 16:
        packing = 0
    chunk = int(rng.choice([0, 0, int(rng.integers(50, n_cols + 1))]))
    return rng, n_cols, lens, dict(segment=segment, precision=precision, packing=packing, ref_chunk=chunk)


@pytest.mark.parametrize("seed", range(24))
def test_random_plan_bit_exact(ctx, seed):
    rng, n_cols, lens, kw = _case(seed)
    ref = rng.standard_normal((2, n_cols)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    flat, offs, lns = D.pack_reads(reads)
    p = ctx.plan(lns, **kw)
    d_q, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lns))
    d_out = D.DeviceArray((len(reads),), D.RESULT_DTYPE)
    p.run(d_q, d_off, d_len, d_out, None)
    res = d_out.to_host()
    p.close()
    qs = [O.f32_to_f16(r) for r in reads]
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    sc, am = O.batch_f16(qs, sh, fused=kw["precision"] == "f16_fused")
    osc = O.f16_to_f32(sc.ravel()).reshape(sc.shape)
    assert np.array_equal(res["score"], osc), (kw, lens)
    assert np.array_equal(res["argmin"], am), (kw, lens)


@pytest.mark.parametrize("seed", range(8))
def test_random_plan_f32_bit_exact(ctx, seed):
    """The binary32 form (the reference's FP32 build arithmetic) over random
    widths and chunkings."""
    rng, n_cols, lens, kw = _case(100 + seed)
    kw = dict(segment=kw["segment"], ref_chunk=kw["ref_chunk"], precision="f32")
    ref = rng.standard_normal((2, n_cols)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    sc, am = ctx.score(reads, **kw)
    osc, oam = O.batch_f32(reads, ref)
    assert np.array_equal(sc, osc), (kw, lens)
    assert np.array_equal(am, oam), (kw, lens)
