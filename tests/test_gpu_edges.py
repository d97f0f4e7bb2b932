"""Edge cases of the scoring path against the oracle: empty batches, tiny
references, queries longer than the reference, a single strand, samples that
overflow binary16, and an empty Read Until batch."""
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


def _oracle_f16(reads, ref):
    qs = [O.f32_to_f16(r) for r in reads]
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(ref.shape[0])])
    sc, am = O.batch_f16(qs, sh)
    return O.f16_to_f32(sc.ravel()).reshape(sc.shape), am


def test_empty_batches(ctx):
    rng = np.random.default_rng(1)
    ctx.upload_reference(rng.standard_normal((2, 500)).astype(np.float32))
    for prec in ("f16", "f16_fused", "f32"):
        sc, am = ctx.score([], precision=prec)
        assert sc.shape == (0, 2) and am.shape == (0, 2)
    p = ctx.plan(np.zeros(0, np.int32))
    assert p.info()["n_items"] == 0
    p.close()
    norm, _ = ctx.normalize([], mode=D.NORM_MEDMAD)
    assert len(norm) == 0
    ru = D.ReadUntil(ctx, max_reads=4, chunk_len=512, max_chunks=2)
    assert len(ru.process([], [], np.zeros((0, 512), np.float32))) == 0
    ru.close()


@pytest.mark.parametrize("n_cols", [1, 2, 63, 64, 65, 130])
def test_tiny_references_and_long_queries(ctx, n_cols):
    rng = np.random.default_rng(n_cols)
    ref = rng.standard_normal((2, n_cols)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (1, 3, 64, 200, 1500)]
    print(f"n_cols={n_cols}", flush=True)
    sc, am = ctx.score(reads, precision="f16")
    osc, oam = _oracle_f16(reads, ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)
    sc, am = ctx.score(reads, precision="f32")
    osc, oam = O.batch_f32(reads, ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)


def test_single_strand_reference(ctx):
    rng = np.random.default_rng(5)
    ref = rng.standard_normal((1, 3000)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (10, 700, 2100)]
    sc, am = ctx.score(reads, precision="f16")
    osc, oam = _oracle_f16(reads, ref)
    assert np.array_equal(sc[:, 0], osc[:, 0]) and np.array_equal(am[:, 0], oam[:, 0])
    assert np.all(np.isinf(sc[:, 1])) and np.all(am[:, 1] == -1)
    sc, am = ctx.score(reads, precision="f32")
    osc, oam = O.batch_f32(reads, ref)
    assert np.array_equal(sc[:, 0], osc[:, 0]) and np.array_equal(am[:, 0], oam[:, 0])


def test_binary16_overflow(ctx):
    """Samples beyond the binary16 range become inf in the binary16 forms:
    every path through that row costs inf, so the read scores inf, exactly
    as the oracle's per-op arithmetic; binary32 keeps a finite score."""
    rng = np.random.default_rng(6)
    ref = rng.standard_normal((2, 2000)).astype(np.float32)
    ctx.upload_reference(ref)
    r = rng.standard_normal(300).astype(np.float32)
    r[150] = 1e6
    ok = rng.standard_normal(300).astype(np.float32)
    sc, am = ctx.score([r, ok], precision="f16")
    osc, oam = _oracle_f16([r, ok], ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)
    assert np.all(np.isinf(sc[0])) and np.all(np.isfinite(sc[1]))
    sc32, _ = ctx.score([r], precision="f32")
    assert np.all(np.isfinite(sc32))
