"""Pins the oracle's clean sDTW recurrence (FP64 / FP32 / binary16 forms)."""
import numpy as np
import pytest

import oracle as O


def _np_matrix(q, s, dt):
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


def _rn16_exact(x):
    """Binary16 round-to-nearest-even of an exact rational x >= 0 (value
    comparison against the neighbouring binary16 values in exact arithmetic,
    so no intermediate rounding)."""
    from fractions import Fraction
    if x >= Fraction(65520):  # halfway between 65504 and the next binade: inf
        return np.float16(np.inf)
    h = np.float16(float(x))  # nearest or a neighbour of it
    with np.errstate(over="ignore"):  # the neighbour above 65504 is inf, dropped below
        cands = {h, np.nextafter(h, np.float16(np.inf)), np.nextafter(h, np.float16(0))}
    cands = [c for c in cands if np.isfinite(c) and c >= 0]
    best = min(abs(Fraction(float(c)) - x) for c in cands)
    tied = [c for c in cands if abs(Fraction(float(c)) - x) == best]
    if len(tied) == 1:
        return tied[0]
    return [c for c in tied if int(c.view(np.uint16)) % 2 == 0][0]


def _fused_matrix_exact(qh, sh):
    """Per-cell restatement of the fused binary16 form with exact rationals:
    t = rn16(s - q); D = rn16(t*t + min(top, diag)), the product and the sum
    exact, one rounding per cell."""
    from fractions import Fraction
    Q, N = len(qh), len(sh)
    D = np.zeros((Q, N), np.float16)
    inf = np.float16(np.inf)
    for r in range(Q):
        for c in range(N):
            t = np.float16(sh[c] - qh[r])  # binary16 subtraction, exactly rounded
            top = np.float16(0) if r == 0 else D[r - 1, c]
            dg = np.float16(0) if r == 0 else (D[r - 1, c - 1] if c > 0 else inf)
            m = min(top, dg)
            if not np.isfinite(m):
                D[r, c] = inf
                continue
            tf = Fraction(float(t))
            D[r, c] = _rn16_exact(tf * tf + Fraction(float(m)))
    return D


def test_rn16_exact_rounding():
    from fractions import Fraction
    for v in (0.0, 1.0, 65504.0, 6e-8, 0.1, 1000.3):
        assert _rn16_exact(Fraction(v)) == np.float16(v)
    one = Fraction(1)
    assert _rn16_exact(one + Fraction(1, 2 ** 11)) == np.float16(1.0)  # tie -> even
    assert _rn16_exact(one + Fraction(3, 2 ** 11)) == np.float16(1 + 2 ** -9)  # tie -> even (up)
    assert _rn16_exact(one + Fraction(1, 2 ** 11) + Fraction(1, 2 ** 40)) == np.float16(1 + 2 ** -10)
    assert np.isinf(_rn16_exact(Fraction(65520)))


@pytest.mark.parametrize("seed", range(3))
def test_f16_fused_per_cell_exact(seed):
    """orc_sdtw_f16_fused equals the exact-rational per-cell restatement of
    D = rn16(t*t + min) (one rounding), on every cell of the last row."""
    rng = np.random.default_rng(200 + seed)
    Q, N = 12, 20
    qh = (rng.standard_normal(Q) * 2).astype(np.float16)
    sh = (rng.standard_normal(N) * 2).astype(np.float16)
    D = _fused_matrix_exact(qh, sh)
    sc, am, row = O.sdtw_f16(qh.view(np.uint16), sh.view(np.uint16), want_row=True, fused=True)
    assert np.array_equal(row, D[-1].view(np.uint16))
    assert am == int(np.argmin(D[-1]))


def test_f16_fused_and_per_op_differ_where_rounding_twice_matters():
    """Two rows, one column: D1 = t1^2 + D0 rounded once (fused) or after
    rounding t1^2 first (per op, the reference's COST_FUNCTION); for these
    values the two differ, and each form equals its own restatement."""
    for qv, sv in (((0.1257, -0.1321), 0.6406), ((1.304, 0.9473), -0.7036)):
        qh = np.array(qv, np.float16)
        sh = np.array([sv], np.float16)
        per_op = _np_matrix(qh, sh, np.float16)
        fused = _fused_matrix_exact(qh, sh)
        assert per_op[-1, 0] != fused[-1, 0]
        a, _ = O.sdtw_f16(qh.view(np.uint16), sh.view(np.uint16))
        b, _ = O.sdtw_f16(qh.view(np.uint16), sh.view(np.uint16), fused=True)
        assert a == per_op[-1].view(np.uint16)[0] and b == fused[-1].view(np.uint16)[0]


def test_f16_fused_close_to_fp64():
    rng = np.random.default_rng(7)
    q = O.f32_to_f16(rng.standard_normal(400))
    s = O.f32_to_f16(rng.standard_normal(2000))
    a, _ = O.sdtw_f16(q, s)
    b, _ = O.sdtw_f16(q, s, fused=True)
    fa, fb = O.f16_to_f32(np.array([a, b], np.uint16))
    ex, _ = O.sdtw_f64(O.f16_to_f32(q), O.f16_to_f32(s))
    assert abs(fa - ex) / ex < 0.05 and abs(fb - ex) / ex < 0.05


def test_baseline_strips_and_isa_builds_agree():
    """The CPU baseline's column strips (a 50,000-column reference is swept
    in strips of 8192) and its x86-64-v4 build give the oracle's scores and
    argmins bit for bit."""
    rng = np.random.default_rng(18)
    strands = rng.standard_normal((2, 50_000)).astype(np.float32)
    qs = [rng.standard_normal(n).astype(np.float32) for n in (1, 40, 301)]
    qs.append(strands[1, 8180:8180 + 30].copy())  # an exact match across a strip edge
    sc, am = O.batch_f32(qs, strands)
    for portable in (True, False):
        sb, ab, _ = O.cpu_baseline_f32(qs, strands, portable=portable)
        assert np.array_equal(sc, sb) and np.array_equal(am, ab), portable
    assert sc[3, 1] == 0 and am[3, 1] == 8180 + 29


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
