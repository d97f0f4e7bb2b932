"""GPU parity of the three-move recurrence (the reference built without
-DNO_REF_DEL, src/include/DTW.cu:23-27: a left move, D[r][c-1], joins top
and diagonal) against the oracle's orc_batch3_*: bit-exact scores and
argmins in binary16 and binary32, for every segment width, multi-band
reads, and per-read widths."""
import numpy as np
import pytest

import dtwax_amd as D
import oracle as O

pytestmark = pytest.mark.gpu

LENS = [1, 2, 63, 64, 65, 100, 511, 512, 513, 1000, 1599, 1600, 2047, 2048, 2049, 2500, 4097]


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def _oracle3_f16(reads, ref):
    qs = [O.f32_to_f16(r) for r in reads]
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(ref.shape[0])])
    sc, am = O.batch3_f16(qs, sh)
    return O.f16_to_f32(sc.ravel()).reshape(sc.shape), am


@pytest.mark.parametrize("segment", [0, 8, 12, 16, 20, 25, 28, 32])
def test_f16_three_moves_bit_exact(ctx, segment):
    rng = np.random.default_rng(300 + segment)
    ref = rng.standard_normal((2, 3001)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in LENS]
    print(f"moves=3 f16 segment={segment}", flush=True)
    sc, am = ctx.score(reads, precision="f16", segment=segment, moves=3)
    osc, oam = _oracle3_f16(reads, ref)
    assert np.array_equal(sc, osc)
    assert np.array_equal(am, oam)
    # right moves only add paths: never above the two-move score
    sc2, _ = ctx.score(reads, precision="f16", segment=segment)
    assert np.all(sc <= sc2)


@pytest.mark.parametrize("nstr", [1, 2])
def test_f32_three_moves_bit_exact(ctx, nstr):
    rng = np.random.default_rng(310 + nstr)
    ref = rng.standard_normal((nstr, 2500)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in LENS[:-2]]
    for seg in (8, 16):
        print(f"moves=3 f32 nstr={nstr} segment={seg}", flush=True)
        sc, am = ctx.score(reads, precision="f32", segment=seg, moves=3)
        osc, oam = O.batch3_f32(reads, ref)
        assert np.array_equal(sc[:, :nstr], osc)
        assert np.array_equal(am[:, :nstr], oam)


def test_three_moves_reference_run_scores_zero(ctx):
    """A query that crosses a run of equal reference values in one sample
    matches it exactly with right moves; two moves give each row one column,
    so they cannot."""
    rng = np.random.default_rng(320)
    ref = O.f16_to_f32(O.f32_to_f16(rng.standard_normal((2, 400)).astype(np.float32)))
    ref = ref.reshape(2, 400).copy()
    ref[0, 100], ref[0, 101:105], ref[0, 105] = 1.0, 0.5, 1.0
    ctx.upload_reference(ref)
    q = np.array([1.0, 0.5, 1.0], np.float32)
    sc3, am3 = ctx.score([q], precision="f16", moves=3)
    sc2, _ = ctx.score([q], precision="f16")
    assert sc3[0, 0] == 0.0 and am3[0, 0] == 105
    assert sc2[0, 0] > 0.0


def test_three_moves_options(ctx):
    rng = np.random.default_rng(321)
    ref = rng.standard_normal((2, 5000)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = np.array([100, 2000], np.int32)
    with pytest.raises(D.DtwaxError, match="whole-strand"):
        ctx.plan(lens, precision="f16", moves=3, ref_chunk=997)
    for prec in ("f16_fused", "f16_renorm"):
        with pytest.raises(D.DtwaxError, match="moves = 3"):
            ctx.plan(lens, precision=prec, moves=3)
    with pytest.raises(D.DtwaxError, match="moves must be"):
        ctx.plan(lens, precision="f16", moves=4)
    # a chunk as long as the reference is a whole strand
    p = ctx.plan(lens, precision="f16", moves=3, ref_chunk=5000)
    p.close()
    reads = [rng.standard_normal(int(n)).astype(np.float32) for n in rng.integers(50, 2000, 24)]
    flat, offs, lns = D.pack_reads(reads)
    d_q, d_off = D.DeviceArray.from_host(flat), D.DeviceArray.from_host(offs)
    best, table = ctx.autotune(d_q, d_off, lns, sample_reads=16, precision="f16", moves=3)
    assert best["ref_chunk"] == 0 and len(table) == 8
    sc, am = ctx.score(reads, precision="f16", moves=3, **best)
    osc, oam = _oracle3_f16(reads, ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)
