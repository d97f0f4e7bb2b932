"""The reference-layout entry points dtwax_distances / dtwax_distances_f32
(mirrors of FullDTW::distances, src/include/DTW.hpp:285-295, in the FP16
and FP32 builds): the reference kernel's own DP with its seam deviations
(compat_quirks 1 on the persistent kernel, 2 on the tile-by-tile kernel),
the clean recurrence (0), the last-row buffer, buffer ownership, and the
size of the deviations' effect on a population at the shipped geometry.

Oracles: the literal emulator of DTW.cu (orc_emulate_reference, pinned in
tests/test_oracle_emulator.py) and the clean recurrence (orc_sdtw_f16/f32)."""
import numpy as np
import pytest

import dtwax_amd as D
import oracle as O
from helpers import (fasta_python, half2_words, load_model_means, query_words_h2, tile_reference,
                     tiled_index, words_to_h2_floats)

pytestmark = pytest.mark.gpu

REF_LEN, TILE = 48128, 832
N_COLS = (REF_LEN // TILE) * TILE  # 47,424 scored columns
REF32_LEN = 2 * REF_LEN            # FP32 build: fwd | rev concatenated
N32_COLS = (REF32_LEN // TILE) * TILE  # 95,680


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def _half_vals(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def _seam_queries(flat, rng, n_random=3):
    """Random queries plus exact reference substrings whose zero-cost path
    crosses a deviating cell: Q1 at row 1024, columns 832k and 832k + 26; Q2
    at rows 962..992 of either 1024-row pass, columns 832k (k >= 2)."""
    starts = [5 * TILE + 26 - 1024, 3 * TILE - 1024, 10 * TILE - 980, 20 * TILE - 1994,
              55 * TILE - 980, 2 * TILE + 26 - 1024]
    return np.concatenate([rng.standard_normal((n_random, 2048)).astype(np.float32),
                           np.stack([flat[c:c + 2048] for c in starts])])


def _run_f16(ctx, flat2, qs, quirks, last_row=False, stream=None):
    n = len(qs)
    d_ref = D.DeviceArray.from_host(tile_reference(half2_words(flat2[0], flat2[1])))
    d_q = D.DeviceArray.from_host(query_words_h2(qs))
    d_out = D.DeviceArray((n,), np.uint32)
    d_lr = None
    if last_row:
        d_lr = D.DeviceArray.from_host(np.full(n * REF_LEN, 0x7e007e00, np.uint32))  # NaN marker
    ctx.distances(d_ref, d_q, d_out, n, compat_quirks=quirks, stream=stream, d_last_row=d_lr)
    D.lib().dtwax_device_sync()
    got = words_to_h2_floats(d_out.to_host())
    if not last_row:
        return got
    return got, words_to_h2_floats(d_lr.to_host()).reshape(n, REF_LEN, 2)


def test_quirk_kernels_equal_emulator_with_last_row(ctx):
    """compat_quirks 1 (deviations applied in the persistent two-column
    sweep) and 2 (the reference's tile-by-tile schedule) both equal the
    literal emulator bit for bit, scores and device_last_row, on random and
    seam-crossing queries; the clean form (0) equals the fixed emulator."""
    rng = np.random.default_rng(301)
    flat = _half_vals(rng.standard_normal((2, REF_LEN)))
    qs = _seam_queries(flat[0], rng)
    want, want_lr = O.emulate_reference(flat, qs, last_row=True)
    clean, clean_lr = O.emulate_reference(flat, qs, fix_q1=True, fix_q2=True, last_row=True)
    assert np.all(clean[3:, 0] == 0) and np.all(want[3:, 0] > 0)  # the deviations matter here
    written = tiled_index(np.arange(N_COLS), REF_LEN)
    for quirks, w, wlr in ((1, want, want_lr), (2, want, want_lr), (0, clean, clean_lr)):
        got, lr = _run_f16(ctx, flat, qs, quirks, last_row=True)
        assert np.array_equal(got, w), (quirks, got, w)
        assert np.array_equal(lr[:, written], wlr[:, written]), quirks
        # entries of no scored column are left as they were
        untouched = np.setdiff1d(np.arange(REF_LEN), written)
        assert untouched.size == REF_LEN - N_COLS and np.all(np.isnan(lr[:, untouched]))


def test_clean_last_row_equals_oracle_row(ctx):
    """compat_quirks 0: the last row written to device_last_row is the clean
    binary16 recurrence's last row (oracle's stored row) at every scored
    column, in the tiled layout."""
    rng = np.random.default_rng(302)
    flat = _half_vals(rng.standard_normal((2, REF_LEN)))
    qs = rng.standard_normal((2, 2048)).astype(np.float32)
    got, lr = _run_f16(ctx, flat, qs, 0, last_row=True)
    idx = tiled_index(np.arange(N_COLS), REF_LEN)
    for r in range(2):
        qh = O.f32_to_f16(qs[r])
        for k in range(2):
            sc, _, row = O.sdtw_f16(qh, O.f32_to_f16(flat[k, :N_COLS]), want_row=True)
            assert np.array_equal(lr[r, idx, k], O.f16_to_f32(row)), (r, k)
            assert got[r, k] == O.f16_to_f32(np.array([sc], np.uint16))[0]


def test_distances_f32_build(ctx):
    """dtwax_distances_f32: the FP32 build's one DP over the concatenated
    fwd | rev squiggle (115 tiles), binary32 per op: compat_quirks 1
    (deviations applied in the persistent binary32 sweep) and 2 (tile by
    tile) equal the emulator in its FP32 mode (scores and last row), 0 the
    clean binary32 recurrence over the 95,680 scored columns."""
    rng = np.random.default_rng(303)
    cat = rng.standard_normal(REF32_LEN).astype(np.float32)
    # random queries; exact substrings whose zero-cost path crosses a Q2 cell
    # (row 980 at column 832*60), the fwd -> rev seam (column 48128), a Q1
    # cell (row 1024 at column 832*7 + 26), and one cell of every other
    # deviating lane: Q1 at 832*3 (lane 32), Q2 rows 1994 (lane 62), 992
    # (lane 31 row 0), 2016 (lane 63 row 0), Q1 at 832*100 + 26
    starts = (60 * TILE - 980, 48000, 7 * TILE + 26 - 1024, 3 * TILE - 1024, 20 * TILE - 1994,
              30 * TILE - 992, 40 * TILE - 2016, 100 * TILE + 26 - 1024)
    qs = np.concatenate([rng.standard_normal((2, 2048)).astype(np.float32),
                         np.stack([cat[c:c + 2048] for c in starts])])
    n = len(qs)
    want, want_lr = O.emulate_reference(cat, qs, ref_len=REF32_LEN, fp16=False, last_row=True)
    written = tiled_index(np.arange(N32_COLS), REF32_LEN)
    d_ref = D.DeviceArray.from_host(tile_reference(cat))
    d_q = D.DeviceArray.from_host(np.concatenate([qs.ravel(), np.zeros(32, np.float32)]))
    d_out = D.DeviceArray((n,), np.float32)
    for quirks in (1, 2, 0):
        d_lr = D.DeviceArray((n * REF32_LEN,), np.float32)
        ctx.distances_f32(d_ref, d_q, d_out, n, compat_quirks=quirks, d_last_row=d_lr)
        D.lib().dtwax_device_sync()
        got = d_out.to_host()
        lr = d_lr.to_host().reshape(n, REF32_LEN)
        if quirks:
            assert np.array_equal(got, want), (got, want)
            assert np.array_equal(lr[:, written], want_lr[:, written])
        else:
            for r in range(n):
                sc, _, row = O.sdtw_f32(qs[r], cat[:N32_COLS], want_row=True)
                assert got[r] == np.float32(sc), (r, got[r], sc)
                assert np.array_equal(lr[r, written], row)
    # the substrings score 0 in the clean DP, even across the strand seam,
    # and above 0 where their path crosses a deviating cell
    assert np.all(got[2:] == 0)
    assert want[3] == 0 and np.all(want[2:][np.arange(len(starts)) != 1] > 0)
    # launches of one and three queries (grids of one and three waves), no last row
    for m in (1, 3):
        d_sub = D.DeviceArray((m,), np.float32)
        ctx.distances_f32(d_ref, d_q, d_sub, m, compat_quirks=1)
        D.lib().dtwax_device_sync()
        assert np.array_equal(d_sub.to_host(), want[:m]), m


def test_distances_keeps_uploaded_reference_two_streams(ctx):
    """dtwax_distances uses its own reference buffer: a dtwax_score batch on
    the uploaded reference and a dtwax_distances batch on another stream of
    the same context, issued together, both equal their oracles; the uploaded
    reference is unchanged afterwards; num_entries may grow and shrink."""
    rng = np.random.default_rng(304)
    lv = rng.standard_normal((2, 20000)).astype(np.float32)
    ctx.upload_reference(lv)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (500, 2048, 3000)]
    flat, offs, lens = D.pack_reads(reads)
    d_rq, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lens))
    d_res = D.DeviceArray((len(reads),), D.RESULT_DTYPE)
    plan = ctx.plan(lens, precision="f16")
    ref_flat = _half_vals(rng.standard_normal((2, REF_LEN)))
    s1, s2 = D.C.c_void_p(), D.C.c_void_p()
    D._check(D.lib().dtwax_stream_create(D.C.byref(s1)))
    D._check(D.lib().dtwax_stream_create(D.C.byref(s2)))
    try:
        d_ref = D.DeviceArray.from_host(tile_reference(half2_words(ref_flat[0], ref_flat[1])))
        for n in (3, 70, 5):
            qs = rng.standard_normal((n, 2048)).astype(np.float32)
            d_q = D.DeviceArray.from_host(query_words_h2(qs))
            d_out = D.DeviceArray((n,), np.uint32)
            plan.run(d_rq, d_off, d_len, d_res, stream=s1.value)
            ctx.distances(d_ref, d_q, d_out, n, compat_quirks=1, stream=s2.value)
            D.lib().dtwax_device_sync()
            got = words_to_h2_floats(d_out.to_host())
            want = O.emulate_reference(ref_flat, qs[:3])
            assert np.array_equal(got[:3], want), n
            res = d_res.to_host()
            osc, oam = O.batch_f16([O.f32_to_f16(r) for r in reads],
                                   np.stack([O.f32_to_f16(lv[0]), O.f32_to_f16(lv[1])]))
            assert np.array_equal(res["score"], O.f16_to_f32(osc.ravel()).reshape(osc.shape)), n
            assert np.array_equal(res["argmin"], oam), n
    finally:
        plan.close()
        D.lib().dtwax_stream_destroy(s1.value)
        D.lib().dtwax_stream_destroy(s2.value)


# Measured on MI355X with the population below (seeded; both GPU forms are
# bit-exact with the emulator): the reference kernel's seam deviations change
# 67 of the 8192 strand scores, by at most 26.125 (a rise of 53.6% on one
# target read), 0.027 on average; no read changes side of the threshold.
POP_CHANGED, POP_MAX_ABS = 67, 26.125


@pytest.fixture(scope="module")
def population(ctx):
    """4096 synthetic reads at the reference's shipped geometry: the lambda
    squiggle of the reference's FASTA and 6-mer model (48,128 k-mers per
    strand), samples [1952, 4000) of 4000-sample reads, half from lambda,
    mean/std normalised with eps 1e-4 -> (levels (2, 48128), queries, is_target)."""
    from dtwax_amd import synth
    _, means = load_model_means()
    seq = fasta_python()
    lv, _ = D.squiggle_build(seq, means, REF_LEN)
    rng = np.random.default_rng(2048)
    reads, is_t, _ = synth.make_reads(rng, seq, means, np.full(4096, 4000), 0.5)
    normed, _ = ctx.normalize([r[1952:4000] for r in reads], mode=D.NORM_MEANSTD, eps=1e-4)
    return lv, np.stack(normed).astype(np.float32), np.asarray(is_t)


def test_shipped_geometry_population(ctx, population):
    """The 4096-read population at the FP16 build's geometry (47,424 scored
    columns per strand): the persistent quirk kernel equals the tile-by-tile
    kernel on every read, the seam deviations move few scores, and the
    decisions at the reference's own threshold (194.07, from its
    mapped/unmapped score tables) and at this population's best threshold
    are identical with and without them."""
    lv, qs, is_t = population
    flat = lv.astype(np.float16).astype(np.float32)
    quirk = _run_f16(ctx, flat, qs, 1)
    tiles = _run_f16(ctx, flat, qs, 2)
    clean = _run_f16(ctx, flat, qs, 0)
    assert np.array_equal(quirk, tiles)
    d = quirk - clean
    changed = int(np.count_nonzero(d))
    print(f"changed {changed}/{d.size} strand scores, max |d| {np.abs(d).max()}, "
          f"mean |d| {np.abs(d).mean():.4f}, max rel {(np.abs(d) / np.maximum(clean, 1e-6)).max():.3f}")
    assert changed == POP_CHANGED and float(np.abs(d).max()) == POP_MAX_ABS
    best_c, best_q = clean.min(axis=1), quirk.min(axis=1)
    assert np.array_equal(best_c < 194.07, best_q < 194.07)
    ths = np.unique(best_c)
    acc = [np.mean((best_c < t) == is_t) for t in ths]
    t_best = ths[int(np.argmax(acc))]
    assert max(acc) > 0.98
    assert np.array_equal(best_c < t_best, best_q < t_best)
    # the reads that moved most, through the literal emulator
    worst = np.argsort(-np.abs(d).max(axis=1))[:3]
    assert np.array_equal(O.emulate_reference(flat, qs[worst]), quirk[worst])


# Measured on MI355X: the seam deviations change 27 of the 4096 FP32-build
# scores, by at most 28.3524 (a rise of 56.2% on one read); no read changes
# side of either threshold.
POP32_CHANGED, POP32_MAX_ABS = 27, 28.3524169921875


def _run_f32(ctx, cat, qs, quirks):
    n = len(qs)
    d_ref = D.DeviceArray.from_host(tile_reference(cat))
    d_q = D.DeviceArray.from_host(np.concatenate([qs.ravel(), np.zeros(32, np.float32)]))
    d_out = D.DeviceArray((n,), np.float32)
    ctx.distances_f32(d_ref, d_q, d_out, n, compat_quirks=quirks)
    D.lib().dtwax_device_sync()
    return d_out.to_host()


def test_fp32_build_population(ctx, population):
    """The same 4096 reads through the FP32 build (the levels fwd | rev as
    96,256 floats, 115 tiles, one DP across the strand seam, binary32 per
    op): the persistent binary32 quirk kernel equals the tile-by-tile kernel
    on every read, and the decisions with and without the seam deviations
    are identical at 194.07 and at this population's best threshold."""
    lv, qs, is_t = population
    cat = np.ascontiguousarray(lv.reshape(-1), np.float32)
    quirk = _run_f32(ctx, cat, qs, 1)
    tiles = _run_f32(ctx, cat, qs, 2)
    clean = _run_f32(ctx, cat, qs, 0)
    assert np.array_equal(quirk, tiles)
    d = quirk - clean
    changed = int(np.count_nonzero(d))
    print(f"FP32 build: changed {changed}/{d.size} scores, max |d| {np.abs(d).max()}, "
          f"mean |d| {np.abs(d).mean():.5f}, max rel {(np.abs(d) / np.maximum(clean, 1e-6)).max():.4f}")
    assert changed == POP32_CHANGED and float(np.abs(d).max()) == POP32_MAX_ABS
    assert np.array_equal(clean < 194.07, quirk < 194.07)
    ths = np.unique(clean)
    acc = [np.mean((clean < t) == is_t) for t in ths]
    t_best = ths[int(np.argmax(acc))]
    assert max(acc) > 0.98
    assert np.array_equal(clean < t_best, quirk < t_best)
    worst = np.argsort(-np.abs(d))[:3]
    assert np.array_equal(O.emulate_reference(cat, qs[worst], ref_len=REF32_LEN, fp16=False), quirk[worst])
