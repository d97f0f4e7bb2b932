"""Pins the oracle's clean sDTW recurrence (FP64 / FP32 / binary16 forms)."""
import numpy as np
import pytest

import oracle as O


def _np_matrix(q, s, dt, fused=False):
    """Per-cell restatement with numpy scalars of dtype dt (independent of the
    C code): D[r][c] = rn(rn(rn(s-q)^2) + min(D[r-1][c], D[r-1][c-1]))."""
    Q, N = len(q), len(s)
    D = np.zeros((Q, N), dt)
    inf = dt(np.inf)
    for r in range(Q):
        for c in range(N):
            t = dt(dt(s[c]) - dt(q[r]))
            top = dt(0) if r == 0 else D[r - 1, c]
            dg = dt(0) if r == 0 else (D[r - 1, c - 1] if c > 0 else inf)
            m = min(top, dg)
            D[r, c] = dt(dt(t * t) + m)
    return D


@pytest.mark.parametrize("seed", range(6))
def test_f64_equals_bruteforce(seed):
    rng = np.random.default_rng(seed)
    Q, N = rng.integers(1, 40), rng.integers(1, 60)
    q, s = rng.standard_normal(Q), rng.standard_normal(N)
    M = O.bruteforce_f64(q, s)
    sc, am, row = O.sdtw_f64(q, s, want_row=True)
    assert np.array_equal(row, M[-1])
    assert sc == M[-1].min() and am == int(np.argmin(M[-1]))


def test_closed_forms():
    rng = np.random.default_rng(3)
    s = rng.standard_normal(500)
    # Q = 1: min_c (q0 - s_c)^2
    sc, am = O.sdtw_f64(np.array([0.37]), s)
    assert sc == np.min((s - 0.37) ** 2) and am == int(np.argmin((s - 0.37) ** 2))
    # exact substring => 0 at its last column
    q = s[123:123 + 77].copy()
    sc, am = O.sdtw_f64(q, s)
    assert sc == 0.0 and am == 123 + 76
    # score is non-decreasing in Q (min over the last row)
    qq = rng.standard_normal(60)
    prev = -1.0
    for k in range(1, 60):
        v, _ = O.sdtw_f64(qq[:k], s)
        assert v >= prev
        prev = v


def test_init_row_extension_equals_from_scratch():
    rng = np.random.default_rng(4)
    s, q = rng.standard_normal(300), rng.standard_normal(90)
    full = O.sdtw_f64(q, s, want_row=True)
    a = O.sdtw_f64(q[:37], s, want_row=True)
    b = O.sdtw_f64(q[37:], s, init_row=a[2], want_row=True)
    assert np.array_equal(full[2], b[2]) and full[:2] == b[:2]
    for fn, dt in ((O.sdtw_f32, np.float32),):
        full = fn(q.astype(dt), s.astype(dt), want_row=True)
        a = fn(q[:37].astype(dt), s.astype(dt), want_row=True)
        b = fn(q[37:].astype(dt), s.astype(dt), init_row=a[2], want_row=True)
        assert np.array_equal(full[2], b[2])
    qh, sh = O.f32_to_f16(q), O.f32_to_f16(s)
    full = O.sdtw_f16(qh, sh, want_row=True)
    a = O.sdtw_f16(qh[:37], sh, want_row=True)
    b = O.sdtw_f16(qh[37:], sh, init_row=a[2], want_row=True)
    assert np.array_equal(full[2], b[2])


@pytest.mark.parametrize("seed", range(3))
def test_f32_and_f16_per_cell(seed):
    rng = np.random.default_rng(100 + seed)
    Q, N = 25, 33
    q = rng.standard_normal(Q).astype(np.float32) * 2
    s = rng.standard_normal(N).astype(np.float32) * 2
    D32 = _np_matrix(q, s, np.float32)
    sc, am, row = O.sdtw_f32(q, s, want_row=True)
    assert np.array_equal(row, D32[-1]) and sc == D32[-1].min()
    qh, sh = q.astype(np.float16), s.astype(np.float16)
    D16 = _np_matrix(qh, sh, np.float16)
    sc, am, row = O.sdtw_f16(qh.view(np.uint16), sh.view(np.uint16), want_row=True)
    assert np.array_equal(row, D16[-1].view(np.uint16))
    assert am == int(np.argmin(D16[-1]))


def test_f16_vector_path_equals_scalar():
    rng = np.random.default_rng(9)
    for Q, N, scale in ((1, 7, 1.0), (37, 1001, 1.0), (300, 3000, 3.0), (200, 777, 1e-3)):
        q = O.f32_to_f16(rng.standard_normal(Q) * scale)
        s = O.f32_to_f16(rng.standard_normal(N) * scale)
        init = O.f32_to_f16(np.abs(rng.standard_normal(N)) * 50)
        for ir in (None, init):
            a = O.sdtw_f16(q, s, init_row=ir, want_row=True)
            b = O.sdtw_f16(q, s, init_row=ir, want_row=True, scalar=True)
            assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2])


def test_f16_fused_close_and_distinct():
    rng = np.random.default_rng(7)
    q = O.f32_to_f16(rng.standard_normal(400))
    s = O.f32_to_f16(rng.standard_normal(2000))
    a, _ = O.sdtw_f16(q, s)
    b, _ = O.sdtw_f16(q, s, fused=True)
    fa, fb = O.f16_to_f32(np.array([a, b], np.uint16))
    ex, _ = O.sdtw_f64(O.f16_to_f32(q), O.f16_to_f32(s))
    assert abs(fa - ex) / ex < 0.05 and abs(fb - ex) / ex < 0.05


def test_batches_and_baseline_agree():
    rng = np.random.default_rng(8)
    strands = rng.standard_normal((2, 700)).astype(np.float32)
    qs = [rng.standard_normal(n).astype(np.float32) for n in (1, 5, 64, 333)]
    sc, am = O.batch_f32(qs, strands)
    sb, ab, used = O.cpu_baseline_f32(qs, strands)
    assert np.array_equal(sc, sb) and np.array_equal(am, ab) and used >= 1
    for i, q in enumerate(qs):
        for k in range(2):
            v, a = O.sdtw_f32(q, strands[k])
            assert v == sc[i, k] and a == am[i, k]
    s64, a64 = O.batch_f64([q.astype(np.float64) for q in qs], strands.astype(np.float64))
    assert np.allclose(s64, sc, rtol=1e-4, atol=1e-5)
