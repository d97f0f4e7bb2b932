"""Pins the literal emulator of the reference kernel (src/include/DTW.cu).

1. With both seam deviations switched off it must equal the clean binary16 /
   FP32 recurrence over the scored columns, bit for bit, for many geometries.
2. With them on, an exact-match query (clean score 0) whose diagonal path
   crosses a seam cell predicted from the code reading must score > 0, and one
   that avoids every predicted seam cell must still score 0.

Predicted seam cells (derived from DTW.cu, geometry W lanes, SEG registers,
TILE = W*SEG, PREFIX rows per band, NUM_WAVES = PREFIX + (TILE-1)//SEG):
  Q1  first row of every band after the first, at columns TILE*k (k >= 1,
      lane 0's diagonal reset to +inf at :315/:442) and TILE*k + SEG (k >= 0,
      lane 1's diagonal still the +inf set at :179-180/:316).
  Q2  tiles k = 2 .. REF_BATCH-1, column TILE*k, band-local rows r with
      r-1 in the entries the middle-tile else-branch (:385-388) rewrites:
      idx = w - (2W-1) for w in [NW_BY_W, NUM_WAVES) with w % W != W-1.
"""
import numpy as np
import pytest

import oracle as O


def _half_ref(rng, ref_len):
    x = rng.standard_normal((2, ref_len)).astype(np.float32)
    return O.f16_to_f32(O.f32_to_f16(x)).reshape(2, ref_len)


@pytest.mark.parametrize("geom", [
    (16, 8, 2, 4, 48), (24, 8, 3, 4, 60), (32, 16, 2, 8, 197), (64, 32, 5, 8, 217),
    (128, 64, 4, 16, 448), (256, 128, 26, 32, 3428), (2048, 1024, 26, 32, 4992)])
def test_fixed_emulator_equals_clean_f16(geom):
    QL, PL, SEG, W, RL = geom
    rng = np.random.default_rng(sum(geom))
    N = (RL // (SEG * W)) * SEG * W
    ref = _half_ref(rng, RL)
    qs = rng.standard_normal((3, QL)).astype(np.float32)
    out = O.emulate_reference(ref, qs, QL, PL, SEG, W, RL, True, True, True)
    raw = O.emulate_reference(ref, qs, QL, PL, SEG, W, RL, True, False, False)
    for r in range(3):
        qh = O.f32_to_f16(qs[r])
        for k in range(2):
            sc, _ = O.sdtw_f16(qh, O.f32_to_f16(ref[k, :N]))
            clean = O.f16_to_f32(np.array([sc], np.uint16))[0]
            assert out[r, k] == clean
            assert raw[r, k] >= clean  # seam deviations only raise cells


@pytest.mark.parametrize("geom", [(64, 32, 5, 8, 217), (96, 32, 3, 8, 24 * 9)])
def test_fixed_emulator_equals_clean_f32(geom):
    QL, PL, SEG, W, RL = geom
    rng = np.random.default_rng(5)
    N = (RL // (SEG * W)) * SEG * W
    ref = rng.standard_normal(RL).astype(np.float32)
    qs = rng.standard_normal((3, QL)).astype(np.float32)
    out = O.emulate_reference(ref, qs, QL, PL, SEG, W, RL, False, True, True)
    for r in range(3):
        sc, _ = O.sdtw_f32(qs[r], ref[:N])
        assert out[r] == sc


def _q2_rows(W, SEG, PL):
    TILE = W * SEG
    nw = PL + (TILE - 1) // SEG
    nwbw = (nw // W) * W
    bad_idx = [w - (2 * W - 1) for w in range(nwbw, nw) if w % W != W - 1]
    return {i + 1 for i in bad_idx}  # local row r uses plc[r-1] as diagonal


def _exact_match_score(W, SEG, PL, QL, RL, c0, fix, seed=0):
    rng = np.random.default_rng(seed)
    ref = _half_ref(rng, RL)
    q = ref[0, c0:c0 + QL].copy()[None, :]
    out = O.emulate_reference(ref, q, QL, PL, SEG, W, RL, True, fix, fix)
    return out[0, 0]


def test_seam_cells_tiny_geometry():
    W, SEG, PL, QL, RL = 4, 2, 16, 32, 48
    TILE = W * SEG
    q2 = _q2_rows(W, SEG, PL)
    assert q2 == {10, 11, 12}
    # Q2: cross column 3*TILE at band-local row 11
    assert _exact_match_score(W, SEG, PL, QL, RL, 3 * TILE - 11, False) > 0
    assert _exact_match_score(W, SEG, PL, QL, RL, 3 * TILE - 11, True) == 0
    # Q1 at TILE*k: cross column 3*TILE at row PL (first row of band 1)
    assert _exact_match_score(W, SEG, PL, QL, RL, 3 * TILE - PL, False) > 0
    assert _exact_match_score(W, SEG, PL, QL, RL, 3 * TILE - PL, True) == 0
    # Q1 at TILE*k + SEG: cross column 2*TILE + SEG at row PL
    assert _exact_match_score(W, SEG, PL, QL, RL, 2 * TILE + SEG - PL, False) > 0
    # control: c0 = 1 crosses seams at rows 7, 15, 23, 31 and row PL at column 17
    assert _exact_match_score(W, SEG, PL, QL, RL, 1, False) == 0


def test_seam_cell_shipped_geometry():
    """Q2 at the shipped geometry: the exact match crosses column 5*832 at
    band-local row 970 (inside 962..992) -> the reference scores it > 0."""
    W, SEG, PL, QL, RL = 32, 26, 1024, 2048, 48128
    assert _q2_rows(W, SEG, PL) == set(range(962, 993))
    c0 = 5 * 832 - 970
    assert _exact_match_score(W, SEG, PL, QL, RL, c0, False) > 0
    # control: crossing at band-local row 500 is untouched
    assert _exact_match_score(W, SEG, PL, QL, RL, 5 * 832 - 500, False) == 0


def test_rejects_geometry_the_reference_cannot_build():
    with pytest.raises(ValueError):
        O.emulate_reference(np.zeros((2, 64), np.float32), np.zeros((1, 16), np.float32),
                            16, 8, 2, 8, 64, True)
