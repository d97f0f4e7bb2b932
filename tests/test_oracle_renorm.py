"""Pins the oracle's renormalised binary16 form (orc_sdtw_f16_renorm).

* With drift 0 and a query of at most one band it is the plain binary16
  recurrence (orc_sdtw_f16) exactly: lowering by 0 changes nothing.
* An independent per-cell restatement with numpy float16 scalars (the
  lane's first row every S rows costing rn(t*t - K) with one rounding,
  taken from exact rationals; a band's top row lowered by its minimum; the
  band-to-band drift rule; the binary32 offsets) gives the same score and
  argmin bit for bit on small inputs with several bands; a plausible slip
  (t*t rounded before the drift comes off, or no top-row lowering) changes
  the result.
* The split schedule (orc_sdtw_f16_renorm_split: 64 x 8-row bands, then
  64 x 16-row bands at twice the drift) against the same restatement.
* It stays within 2% of FP64 where plain binary16 drifts several percent.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

H = np.float16


def _bands(Q, S, split):
    """(rows per lane, first read row of the band) for every band: bands of
    64*S rows, front-padded; split: the rows after the first 513..1536 in
    bands of 64*2S rows (reads over 1536 rows)."""
    rest = 1024 * ((Q - 513) // 1024) if (split and Q > 1536) else 0
    Q1 = Q - rest
    nb1 = -(-Q1 // (64 * S))
    P = nb1 * 64 * S - Q1
    out = [(S, b * 64 * S - P) for b in range(nb1)]
    out += [(2 * S, Q1 + b * 128 * S) for b in range(rest // (128 * S))]
    return out


def _rn16(x):
    """Exact rational x rounded to binary16 (nearest, ties to even)."""
    h = np.float16(float(x))
    if not np.isfinite(h):
        return h
    cands = [np.nextafter(h, H(-np.inf)), h, np.nextafter(h, H(np.inf))]
    return min(cands, key=lambda v: (abs(Fraction(float(v)) - x), int(np.array(v).view(np.uint16)) & 1))


def _np_renorm(qh, sh, S, drift=0.01, fold=True, top_min=True, split=False, double=True):
    Q, N = len(qh), len(sh)
    bands = _bands(Q, S, split)
    nb = len(bands)
    a = np.zeros(N, H)  # row -1
    Kb, prev_min, off = H(np.float32(drift) * np.float32(S)), H(0), np.float32(0)
    inf = H(np.inf)
    for band, (Sb, rb) in enumerate(bands):
        if double and band > 0 and Sb != bands[band - 1][0]:
            Kb = H(Kb + Kb)  # twice the rows per lane
        low = prev_min if (band > 0 and top_min) else H(0)
        if band > 0:  # the top row, staged lowered by its minimum
            a = np.array([H(v - low) for v in a], H)
        for i in range(64 * Sb):
            r = rb + i
            new = np.empty(N, H)
            for c in range(N):
                t = H(sh[c] - qh[r]) if r >= 0 else H(0)
                if i % Sb == 0:  # the lane's first row: rn(t*t - Kb), one rounding
                    t2 = _rn16(Fraction(float(t)) ** 2 - Fraction(float(Kb))) if fold else H(H(t * t) - Kb)
                else:
                    t2 = H(t * t)
                dg = a[c - 1] if c > 0 else inf
                new[c] = H(t2 + min(a[c], dg))
            a = new
        off = np.float32(off + np.float32(np.float32(64) * np.float32(Kb) + np.float32(low)))
        if band + 1 < nb:
            m = a.min()
            prev_min = m
            Kb = H(Kb + H(m * H(1 / 64)))
    c = int(np.argmin(a))
    return np.float32(np.float32(a[c]) + off), c


def test_drift_zero_one_band_is_plain_f16():
    rng = np.random.default_rng(0)
    for S, Q in ((8, 512), (8, 300), (16, 1000), (32, 1)):
        q = O.f32_to_f16(rng.standard_normal(Q))
        s = O.f32_to_f16(rng.standard_normal(3000))
        sc, am = O.sdtw_f16_renorm(q, s, S, drift=0.0)
        p, pa = O.sdtw_f16(q, s)
        assert sc == O.f16_to_f32(np.array([p], np.uint16))[0] and am == pa, (S, Q)


# seeds on which each slip below changes the score (rounding t*t before the
# drift comes off changes few cells, so not every input shows it)
@pytest.mark.parametrize("S,Q,N,seed", [(8, 700, 37, 17), (8, 1100, 23, 1), (12, 900, 31, 11)])
def test_per_cell_numpy_restatement(S, Q, N, seed):
    rng = np.random.default_rng(seed)
    qh = (rng.standard_normal(Q) * 1.5).astype(H)
    sh = (rng.standard_normal(N) * 1.5).astype(H)
    want, wam = _np_renorm(qh, sh, S)
    got, gam = O.sdtw_f16_renorm(qh.view(np.uint16), sh.view(np.uint16), S)
    assert np.float32(got) == want and gam == wam, (got, want, gam, wam)
    # plausible slips give a different number: rounding t*t before the
    # drift is taken off, or no lowering of a band's top row
    slips = [_np_renorm(qh, sh, S, fold=False)[0], _np_renorm(qh, sh, S, top_min=False)[0]]
    assert all(sl != want for sl in slips), (slips, want)


def test_split_schedule_per_cell():
    """The split schedule (orc_sdtw_f16_renorm_split): 1600 rows are two
    bands of 64 x 8 rows (448 virtual rows) and one of 64 x 16, the drift
    doubled at the switch; leaving out the doubling or the switch changes
    the result."""
    assert [O.renorm_split_rows(q) for q in (1000, 1536, 1537, 1600, 2000, 2561, 4000)] == \
        [0, 0, 1024, 1024, 1024, 2048, 3072]
    assert _bands(1600, 8, True) == [(8, -448), (8, 64), (16, 576)]
    rng = np.random.default_rng(7)
    qh = (rng.standard_normal(1600) * 1.5).astype(H)
    sh = (rng.standard_normal(9) * 1.5).astype(H)
    want, wam = _np_renorm(qh, sh, 8, split=True)
    got, gam = O.sdtw_f16_renorm_split(qh.view(np.uint16), sh.view(np.uint16))
    assert np.float32(got) == want and gam == wam, (got, want, gam, wam)
    slips = [_np_renorm(qh, sh, 8, split=True, double=False)[0], _np_renorm(qh, sh, 8)[0]]
    assert all(sl != want for sl in slips), (slips, want)
    # up to 1536 rows it is the plain schedule at S = 8
    q2 = O.f32_to_f16(rng.standard_normal(1536))
    s2 = O.f32_to_f16(rng.standard_normal(700))
    assert O.sdtw_f16_renorm_split(q2, s2) == O.sdtw_f16_renorm(q2, s2, 8)


def test_batch_equals_single_and_tracks_fp64():
    rng = np.random.default_rng(5)
    ref = rng.standard_normal((2, 4000)).astype(np.float32)
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    qs = [rng.standard_normal(n).astype(np.float32) for n in (100, 2000, 5000)]
    qs.append(ref[1, 500:3500] + rng.normal(0, 0.3, 3000).astype(np.float32))
    qh = [O.f32_to_f16(q) for q in qs]
    sc, am = O.batch_f16_renorm(qh, sh, 8)
    sc0, am0 = O.batch_f16_renorm(qh, sh, 0)  # the split schedule
    for i, q in enumerate(qh):
        for k in range(2):
            v, a = O.sdtw_f16_renorm(q, sh[k], 8)
            assert sc[i, k] == np.float32(v) and am[i, k] == a
            v, a = O.sdtw_f16_renorm_split(q, sh[k])
            assert sc0[i, k] == np.float32(v) and am0[i, k] == a
    for i, q in enumerate(qs):
        f64 = min(O.sdtw_f64(O.f16_to_f32(qh[i]).astype(np.float64), O.f16_to_f32(sh[k]).astype(np.float64))[0]
                  for k in range(2))
        f16 = min(O.f16_to_f32(np.array([O.sdtw_f16(qh[i], sh[k])[0]], np.uint16))[0] for k in range(2))
        for rn in (sc[i].min(), sc0[i].min()):
            assert abs(rn - f64) / f64 < 0.02, (i, rn, f64)
            if len(q) >= 2000:
                assert abs(rn - f64) < abs(f16 - f64), (i, rn, f16, f64)
