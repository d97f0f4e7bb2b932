"""GPU parity of packed plans: several reads' partial bands sharing one band.

A packed read's rows are a partial band (its length modulo 64 x S rows,
front-padded to a lane) that shares a band with other reads, then its full
bands.  The arithmetic of every cell is unchanged, so scores and argmins
stay bit-exact with the oracle, for the per-op and the fused binary16 forms.
"""
import numpy as np
import pytest

import dtwax_amd as D
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def _oracle(reads, ref, fused=False):
    qs = [O.f32_to_f16(r) for r in reads]
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(ref.shape[0])])
    sc, am = O.batch_f16(qs, sh, fused=fused)
    return O.f16_to_f32(sc.ravel()).reshape(sc.shape), am


def _run(ctx, reads, **kw):
    flat, offs, lns = D.pack_reads(reads)
    p = ctx.plan(lns, **kw)
    info = p.info()
    d_q, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lns))
    d_out = D.DeviceArray((len(reads),), D.RESULT_DTYPE)
    p.run(d_q, d_off, d_len, d_out, None)
    res = d_out.to_host()
    p.close()
    return res["score"], res["argmin"], info


# lane counts of the partial bands at S = 32: 4..64 lanes single-band reads,
# continuing reads with 5, 13, 29 and 63-lane partial bands, and reads that
# stay unpacked (partial band under 4 lanes: 1, 63, 2049, 4097 samples)
LENS32 = [1, 63, 97, 128, 129, 300, 511, 700, 1000, 1500, 2000, 2048, 2049, 2200, 2460, 2950,
          4097, 5000, 6100, 1024, 160, 257]


@pytest.mark.parametrize("precision", ["f16", "f16_fused"])
@pytest.mark.parametrize("chunk", [0, 700])
def test_packed_bit_exact(ctx, precision, chunk):
    rng = np.random.default_rng(5 + chunk)
    ref = rng.standard_normal((2, 2500)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in LENS32]
    sc, am, info = _run(ctx, reads, precision=precision, packing=1, ref_chunk=chunk)
    assert info["packs"] > 0 and info["packed_reads"] == len(LENS32) - 4
    osc, oam = _oracle(reads, ref, fused=precision == "f16_fused")
    assert np.array_equal(sc, osc)
    assert np.array_equal(am, oam)


def test_packed_width16_and_scratch_rows(ctx):
    """S = 16 packs; partial bands of many continuing reads (two scratch
    rows per wave) and flush periods 4, 8, 16 and 64."""
    rng = np.random.default_rng(17)
    ref = rng.standard_normal((2, 1800)).astype(np.float32)
    ctx.upload_reference(ref)
    # at S = 16 a band is 1024 rows; partial lanes 4, 5, 8, 9, 16, 17, 33, 64
    lens = [64, 80, 128, 144, 256, 272, 528, 1024] + [1024 + 200] * 9 + [3 * 1024 + 700] * 3
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    sc, am, info = _run(ctx, reads, precision="f16", packing=1, segment=16, ref_chunk=600)
    assert info["packed_reads"] == len(lens)
    osc, oam = _oracle(reads, ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)


@pytest.mark.parametrize("segment,precision", [(20, "f16"), (25, "f16"), (28, "f16_fused"), (25, "f16_fused")])
def test_packed_other_widths(ctx, segment, precision):
    """Packing at S = 20, 25 and 28: 2000-sample reads at S = 25 are one full
    band and a 16-lane partial band, four partials to a shared band."""
    rng = np.random.default_rng(segment)
    ref = rng.standard_normal((2, 1300)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = [2000] * 6 + [300, 777, 1280, 1281, 2600, 4100, 150]
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    sc, am, info = _run(ctx, reads, precision=precision, packing=1, segment=segment, ref_chunk=500)
    assert info["packs"] > 0 and info["segment"] in (0, segment)
    osc, oam = _oracle(reads, ref, fused=precision == "f16_fused")
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)


def test_automatic_packing_uneven(ctx):
    """Automatic plans pack uneven reads that gain, leave a uniform batch
    unpacked, and every layout scores the same."""
    rng = np.random.default_rng(23)
    ref = rng.standard_normal((2, 1500)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = rng.integers(500, 8001, 48)
    reads = [rng.standard_normal(int(n)).astype(np.float32) for n in lens]
    a_sc, a_am, a_info = _run(ctx, reads, precision="f16")
    n_sc, n_am, n_info = _run(ctx, reads, precision="f16", packing=-1)
    assert a_info["packs"] > 0 and n_info["packs"] == 0
    assert a_info["computed"] < n_info["computed"]
    assert np.array_equal(a_sc, n_sc) and np.array_equal(a_am, n_am)
    osc, oam = _oracle(reads, ref)
    assert np.array_equal(a_sc, osc) and np.array_equal(a_am, oam)
    u = [rng.standard_normal(2000).astype(np.float32) for _ in range(8)]
    _, _, u_info = _run(ctx, u, precision="f16")
    assert u_info["packs"] == 0  # 63 of 64 lanes: nothing to gain


def test_packing_option_errors(ctx):
    ref = np.zeros((2, 100), np.float32)
    ctx.upload_reference(ref)
    for kw, msg in [(dict(packing=2), "packing must"),
                    (dict(packing=1, precision="f32"), "packing needs"),
                    (dict(packing=1, precision="f16_renorm"), "packing needs"),
                    (dict(packing=1, moves=3), "packing needs"),
                    (dict(packing=1, segment=8), "packing needs")]:
        with pytest.raises(D.DtwaxError, match=msg):
            ctx.plan(np.array([100], np.int32), **kw)
