"""Pins the oracle's binary16 arithmetic against numpy and exact integer math."""
from fractions import Fraction

import numpy as np

import oracle as O


def test_f32_to_f16_matches_numpy_random():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2**32, size=2_000_000, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    f = f[np.isfinite(f)]
    with np.errstate(over="ignore"):
        ref = f.astype(np.float16).view(np.uint16)
    assert np.array_equal(O.f32_to_f16(f), ref)


def test_f32_to_f16_boundaries():
    h = np.arange(65536, dtype=np.uint16)
    hf = h.view(np.float16).astype(np.float32)
    fin = hf[np.isfinite(hf)]
    fin = np.unique(fin)
    mids = ((fin[:-1].astype(np.float64) + fin[1:].astype(np.float64)) / 2).astype(np.float32)
    cands = [fin, np.nextafter(fin, np.float32(np.inf)), np.nextafter(fin, np.float32(-np.inf)),
             mids, np.nextafter(mids, np.float32(np.inf)), np.nextafter(mids, np.float32(-np.inf)),
             np.array([65504, 65519.99, 65520, 65536, 1e30, 2.0**-25, 2.0**-24, 2.0**-14,
                       -2.0**-25, 0.0, -0.0, np.inf, -np.inf], np.float32)]
    for arr in cands:
        with np.errstate(over="ignore"):
            assert np.array_equal(O.f32_to_f16(arr), arr.astype(np.float16).view(np.uint16))


def test_f16_to_f32_all_encodings():
    h = np.arange(65536, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    mine = O.f16_to_f32(h)
    fin = np.isfinite(ref)
    assert np.array_equal(mine[fin].view(np.uint32), ref[fin].view(np.uint32))
    assert np.all(np.isnan(mine[np.isnan(ref)]))


def test_ops_match_numpy_float16():
    rng = np.random.default_rng(12)
    a = (rng.standard_normal(20000) * 30).astype(np.float16)
    b = (rng.standard_normal(20000) * 3).astype(np.float16)
    L = O.lib()
    au, bu = a.view(np.uint16), b.view(np.uint16)
    for i in range(0, 20000, 7):
        assert L.orc_h_add(int(au[i]), int(bu[i])) == np.float16(a[i] + b[i]).view(np.uint16)
        assert L.orc_h_sub(int(au[i]), int(bu[i])) == np.float16(a[i] - b[i]).view(np.uint16)
        assert L.orc_h_mul(int(au[i]), int(bu[i])) == np.float16(a[i] * b[i]).view(np.uint16)


def _exact_round_f16(x: Fraction) -> int:
    """Independent round-to-nearest-even of an exact rational to binary16 bits."""
    sign = 0x8000 if x < 0 else 0
    x = abs(x)
    if x == 0:
        return sign
    if x >= 65520:
        return sign | 0x7c00
    if x < Fraction(1, 2**14):
        field, quantum = 0, Fraction(1, 2**24)
    else:
        field = next(f for f in range(1, 31)
                     if Fraction(2) ** (f - 15) <= x < Fraction(2) ** (f - 14))
        quantum = Fraction(2) ** (field - 25)
    q = x / quantum
    n = q.numerator // q.denominator
    rem = q - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    if field == 0:
        return sign | n  # n == 1024 encodes the smallest normal
    if n == 2048:
        field, n = field + 1, 1024
    if field >= 31:
        return sign | 0x7c00
    return sign | (field << 10) | (n - 1024)


def test_fused_fma_exact():
    rng = np.random.default_rng(13)
    L = O.lib()
    vals = np.concatenate([(rng.standard_normal(3000) * s).astype(np.float16)
                           for s in (1e-6, 1e-3, 0.1, 1, 10, 300)])
    vals = vals[np.isfinite(vals)]
    for i in range(0, vals.size - 2, 3):
        a, b, c = vals[i], vals[i + 1], abs(vals[i + 2])
        ex = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
        want = _exact_round_f16(ex)
        got = L.orc_h_fma(int(a.view(np.uint16)), int(b.view(np.uint16)), int(c.view(np.uint16)))
        assert got == want, (float(a), float(b), float(c))
