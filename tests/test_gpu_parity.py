"""GPU parity: the HIP kernels through the C ABI against the CPU oracle.

Bar: bit-exact scores and argmins for every arithmetic form the oracle
restates (binary16 per-op, binary16 fused, binary32), for every segment
width and reference chunking; FP64-relative tolerance for the renormalised
form (its offsets change the rounding by design).
"""
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


def _half_vals(x):
    return O.f16_to_f32(O.f32_to_f16(x)).reshape(np.shape(x))


def test_lane_primitives():
    shr, shl = D.selftest_lanes()
    assert list(shr) == [-1] + list(range(63)), "DPP wave_shr:1 must move lane L-1 -> L"
    assert list(shl) == list(range(1, 64)) + [-1], "DPP wave_shl:1 must move lane L+1 -> L"


def _oracle_f16(reads, ref, fused=False):
    qs = [O.f32_to_f16(r) for r in reads]
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(ref.shape[0])])
    sc, am = O.batch_f16(qs, sh, fused=fused)
    return O.f16_to_f32(sc.ravel()).reshape(sc.shape), am


@pytest.mark.parametrize("segment", [8, 12, 16, 20, 25, 28, 32])
def test_f16_bit_exact(ctx, segment):
    rng = np.random.default_rng(segment)
    ref = rng.standard_normal((2, 3001)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in LENS]
    sc, am = ctx.score(reads, precision="f16", segment=segment)
    osc, oam = _oracle_f16(reads, ref)
    assert np.array_equal(sc, osc)
    assert np.array_equal(am, oam)


@pytest.mark.parametrize("segment", [8, 16, 20, 32])
def test_f16_fused_bit_exact(ctx, segment):
    rng = np.random.default_rng(41)
    ref = rng.standard_normal((2, 900)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (1, 64, 130, 700, 2100)]
    sc, am = ctx.score(reads, precision="f16_fused", segment=segment)
    osc, oam = _oracle_f16(reads, ref, fused=True)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)


@pytest.mark.parametrize("nstr", [1, 2])
def test_f32_bit_exact(ctx, nstr):
    rng = np.random.default_rng(50 + nstr)
    ref = rng.standard_normal((nstr, 2500)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in LENS[:-2]]
    for seg in (8, 12, 16, 20, 25, 28, 32):  # 25..32: row costs looked ahead across steps
        sc, am = ctx.score(reads, precision="f32", segment=seg)
        osc, oam = O.batch_f32(reads, ref)
        assert np.array_equal(sc[:, :nstr], osc)
        assert np.array_equal(am[:, :nstr], oam)
        if nstr == 1:
            assert np.all(am[:, 1] == -1)


@pytest.mark.parametrize("prec", ["f16", "f32"])
def test_reference_chunking_is_exact(ctx, prec):
    """Chunks overlapping by (Q-1) columns give the unchunked answer exactly."""
    rng = np.random.default_rng(61)
    ref = rng.standard_normal((2, 20000)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (5, 300, 1100, 2100)]
    full = ctx.score(reads, precision=prec, ref_chunk=0)
    for chunk in (997, 4096, 7000):
        part = ctx.score(reads, precision=prec, ref_chunk=chunk)
        assert np.array_equal(full[0], part[0]) and np.array_equal(full[1], part[1])


def test_renorm_chunked_items(ctx):
    """DTWAX_F16_RENORM with the reference in chunks: every item takes its
    bands' row minima over its own columns, so its subtracted constants, and
    with them the binary16 rounding, depend on the chunking; each chunked
    run equals the oracle's renormalised form on each chunk's columns (x to
    c1) taken separately and combined by min, bit for bit, and stays within
    rounding of the whole-strand run.  The 2100-sample read runs the split
    schedule, its band scratch indexed from the item's first column."""
    rng = np.random.default_rng(61)
    N = 20000
    ref = rng.standard_normal((2, N)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = (5, 300, 1100, 2100)
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    reads.append(ref[0, 11000:13000] + rng.normal(0, 0.1, 2000).astype(np.float32))  # crosses 12288
    full = ctx.score(reads, precision="f16_renorm", ref_chunk=0)
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    qh = [O.f32_to_f16(r) for r in reads]
    for chunk in (4096, 7000):
        sc, am = ctx.score(reads, precision="f16_renorm", ref_chunk=chunk)
        for i, q in enumerate(qh):
            for k in range(2):
                best = (np.float32(np.inf), -1)
                for c0 in range(0, N, chunk):
                    c1 = min(N, c0 + chunk)
                    x = max(0, c0 - (len(q) - 1))
                    # the item's DP over [x, c1) with its own renormalisation;
                    # its minimum counts over [c0, c1) only
                    row, off = O.sdtw_f16_renorm_row(q, sh[k, x:c1], S=8, split=len(q) > 1536)
                    j = int(np.argmin(row[c0 - x:]))
                    v = np.float32(row[c0 - x + j] + off)
                    if v < best[0]:
                        best = (v, c0 + j)
                assert sc[i, k] == best[0] and am[i, k] == best[1], (chunk, i, k, sc[i, k], best)
        rel = np.abs(sc - full[0]) / np.abs(full[0])
        assert rel.max() < 5e-3, (chunk, rel.max())


def test_exact_substring_scores_zero(ctx):
    rng = np.random.default_rng(62)
    ref = _half_vals(rng.standard_normal((2, 5000)).astype(np.float32)).reshape(2, 5000)
    ctx.upload_reference(ref)
    q = ref[1, 1234:1234 + 1777].copy()
    sc, am = ctx.score([q], precision="f16")
    assert sc[0, 1] == 0.0 and am[0, 1] == 1234 + 1776


def test_renorm_tracks_fp64(ctx):
    """At Q = 8000 on random (non-matching) data plain binary16 drifts ~22%
    low; the renormalised form must stay within 2% of FP64 and cut the plain
    form's error at least fivefold on every read."""
    rng = np.random.default_rng(63)
    ref = rng.standard_normal((2, 6000)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(8000).astype(np.float32) for _ in range(6)]
    ex, _ = O.batch_f64([r.astype(np.float64) for r in reads], _half_vals(ref).reshape(2, -1).astype(np.float64))
    f16, _ = ctx.score(reads, precision="f16")
    ren, _ = ctx.score(reads, precision="f16_renorm")
    e16 = np.abs(f16 - ex) / ex
    ere = np.abs(ren - ex) / ex
    print("f16 rel err", e16.round(4).tolist(), "renorm rel err", ere.round(4).tolist())
    assert np.all(ere < 0.02), ere
    assert np.all(ere < 0.2 * e16), (ere, e16)


def test_normalize_medmad_exact(ctx):
    rng = np.random.default_rng(64)
    reads = [(rng.standard_normal(n) * 7 + 90).astype(np.float32) for n in (1, 2, 3, 1000, 2000, 4001)]
    out, st = ctx.normalize(reads, mode=D.NORM_MEDMAD)
    for r, y, (c, s) in zip(reads, out, st):
        want, med, mad = O.norm_medmad(r)
        assert c == med and np.isclose(s, 1.4826 * mad, rtol=0, atol=0)
        assert np.array_equal(y, want)


def test_normalize_medmad_ties(ctx):
    """Integer ADC-like samples (many repeated values), even and odd counts,
    constant reads (MAD 0) and two-valued reads: the middle order statistics
    and the MAD stay exact when the middle keys repeat or differ."""
    rng = np.random.default_rng(67)
    reads = [rng.integers(400, 520, n).astype(np.float32) for n in (2, 4, 1000, 2000, 2001, 4096)]
    reads += [np.full(10, 7.0, np.float32), np.full(11, -3.0, np.float32),
              np.array([1, 1, 2, 2], np.float32), np.array([5, 5, 5, 9, 9, 9], np.float32),
              np.repeat(np.float32([3, 8]), [999, 1001]), np.float32([0.0, -0.0, 0.0, -0.0])]
    out, st = ctx.normalize(reads, mode=D.NORM_MEDMAD)
    for r, y, (c, s) in zip(reads, out, st):
        want, med, mad = O.norm_medmad(r)
        assert c == med and s == 1.4826 * mad, (r[:8], c, med, s, mad)
        assert np.array_equal(y, want)


def test_normalize_meanstd(ctx):
    rng = np.random.default_rng(65)
    reads = [(rng.standard_normal(n) * 13 + 500).astype(np.float32) for n in (2, 1000, 2048)]
    out, st = ctx.normalize(reads, mode=D.NORM_MEANSTD, eps=1e-4)
    for r, y in zip(reads, out):
        want, m, s = O.norm_meanstd(r, 1e-4)
        # FP64 sums in a different order: at most one float ulp apart
        assert np.all(np.abs(y - want) <= np.spacing(np.abs(want)))


def test_distances_mirror_clean(ctx):
    """dtwax_distances (FullDTW::distances mirror, reference layouts) with the
    seam deviations off equals the fixed emulator, i.e. the clean binary16
    recurrence over the 57 scored tiles."""
    rng = np.random.default_rng(66)
    REF_LEN, TILE, SEG, W = 48128, 832, 26, 32
    flat = _half_vals(rng.standard_normal((2, REF_LEN)).astype(np.float32)).reshape(2, REF_LEN)
    # tiled layout h[l*TILE + i*W + j] = flat[l*TILE + j*SEG + i] (main.cu:92-109)
    tiled = np.zeros((REF_LEN, 2), np.float16)
    for l in range(REF_LEN // TILE):
        for i in range(SEG):
            for k in range(2):
                tiled[l * TILE + i * W:l * TILE + i * W + W, k] = flat[k, l * TILE + np.arange(W) * SEG + i]
    n = 3
    qs = rng.standard_normal((n, 2048)).astype(np.float32)
    qh2 = np.zeros((n * 2048 + 32, 2), np.float16)
    qh2[: n * 2048, 0] = qs.ravel()
    qh2[: n * 2048, 1] = qs.ravel()
    d_ref = D.DeviceArray.from_host(tiled.view(np.uint32).ravel())
    d_q = D.DeviceArray.from_host(qh2.view(np.uint32).ravel())
    d_out = D.DeviceArray((n,), np.uint32)
    ctx.distances(d_ref, d_q, d_out, n, compat_quirks=0)
    D.lib().dtwax_device_sync()
    got = d_out.to_host().view(np.float16).reshape(n, 2).astype(np.float32)
    want = O.emulate_reference(flat, qs, fix_q1=True, fix_q2=True)
    assert np.array_equal(got, want)


def test_distances_mirror_quirks(ctx):
    """dtwax_distances with compat_quirks=1 reproduces the reference kernel
    bit for bit, seam deviations included: equal to the literal emulator on
    random queries and on exact reference substrings whose zero-cost path
    crosses a deviating cell (Q1 at row 1024, columns 832k and 832k + 26;
    Q2 at rows 962..992 of either band, columns 832k), where the clean
    recurrence would give 0."""
    rng = np.random.default_rng(67)
    REF_LEN, TILE, SEG, W = 48128, 832, 26, 32
    flat = _half_vals(rng.standard_normal((2, REF_LEN)).astype(np.float32)).reshape(2, REF_LEN)
    tiled = np.zeros((REF_LEN, 2), np.float16)
    for l in range(REF_LEN // TILE):
        for i in range(SEG):
            for k in range(2):
                tiled[l * TILE + i * W:l * TILE + i * W + W, k] = flat[k, l * TILE + np.arange(W) * SEG + i]
    starts = [5 * TILE + 26 - 1024,   # Q1 at 832k + 26
              3 * TILE - 1024,        # Q1 at 832k
              10 * TILE - 980,        # Q2, band 0
              20 * TILE - 1994]       # Q2, band 1
    qs = np.concatenate([rng.standard_normal((3, 2048)).astype(np.float32),
                         np.stack([flat[0, c:c + 2048] for c in starts])])
    n = len(qs)
    qh2 = np.zeros((n * 2048 + 32, 2), np.float16)
    qh2[: n * 2048, 0] = qs.ravel()
    qh2[: n * 2048, 1] = qs.ravel()
    d_ref = D.DeviceArray.from_host(tiled.view(np.uint32).ravel())
    d_q = D.DeviceArray.from_host(qh2.view(np.uint32).ravel())
    d_out = D.DeviceArray((n,), np.uint32)
    ctx.distances(d_ref, d_q, d_out, n, compat_quirks=1)
    D.lib().dtwax_device_sync()
    got = d_out.to_host().view(np.float16).reshape(n, 2).astype(np.float32)
    want = O.emulate_reference(flat, qs)
    clean = O.emulate_reference(flat, qs, fix_q1=True, fix_q2=True)
    assert np.all(clean[3:, 0] == 0) and np.all(want[3:, 0] > 0)  # the deviations matter here
    assert np.array_equal(got, want), (got, want)


def test_c2_subset_decisions(ctx):
    """64 reads of the 48.5 kb synthetic workload (3000 raw -> trim 1000 ->
    2000, median/MAD on the GPU): binary16 scores and argmins bit-exact with
    the oracle, hence identical decisions at any threshold."""
    from dtwax_amd import synth
    w = synth.config_workload(seed=2, ref_bp=48_500, n_reads=64, raw_len=3000, target_frac=0.5)
    reads = synth.trim_reads(w.reads, 1000)
    norm, _ = ctx.normalize(reads, mode=D.NORM_MEDMAD)
    ctx.upload_reference(w.ref_levels)
    sc, am = ctx.score(norm, precision="f16")
    osc, oam = _oracle_f16(norm, w.ref_levels)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)
    best = sc.min(axis=1)
    t = np.median(best)
    assert np.array_equal(best < t, osc.min(axis=1) < t)
    # the synthetic targets and non-targets separate at the best threshold
    acc = max(np.mean((best < th) == w.is_target) for th in np.unique(best))
    assert acc >= 0.9, acc


def test_mixed_segment_plan_exact(ctx):
    """Automatic widths split uneven reads over several persistent launches
    (one per width); scores and argmins stay bit-exact with the oracle."""
    rng = np.random.default_rng(91)
    ref = rng.standard_normal((2, 4001)).astype(np.float32)
    ctx.upload_reference(ref)
    # lengths whose least-padding widths are 8, 12, 16, 20, 25, 28 and 32;
    # 64-column pieces give every width group enough work items to keep its
    # own launch
    lens = ([500] * 140 + [760] * 140 + [1000] * 140 + [1270] * 140 + [1600] * 140 +
            [1790] * 140 + [2048] * 140 + [777, 3001])
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    flat, offs, lns = D.pack_reads(reads)
    p = ctx.plan(lns, precision="f16", segment=0, ref_chunk=64)
    assert p.info()["segment"] == 0  # several widths in this plan
    d_q, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lns))
    d_out = D.DeviceArray((len(reads),), D.RESULT_DTYPE)
    p.run(d_q, d_off, d_len, d_out, None)
    res = d_out.to_host()
    p.close()
    osc, oam = _oracle_f16(reads, ref)
    assert np.array_equal(res["score"], osc) and np.array_equal(res["argmin"], oam)


def test_autotune_choice_cached_and_exact(ctx, tmp_path):
    """The autotuner times every candidate on a sample, caches its choice
    (memory and file), and the chosen plan scores exactly like any other."""
    rng = np.random.default_rng(77)
    ref = rng.standard_normal((2, 20000)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = rng.integers(100, 3000, 40)
    reads = [rng.standard_normal(int(n)).astype(np.float32) for n in lens]
    flat, offs, lns = D.pack_reads(reads)
    d_q, d_off = D.DeviceArray.from_host(flat), D.DeviceArray.from_host(offs)
    cache = tmp_path / "tune.txt"
    best, table = ctx.autotune(d_q, d_off, lns, sample_reads=16, sample_cols=8192, cache_path=cache)
    assert len(table) == 8  # widths 0 (per read), 8, 12, 16, 20, 25, 28, 32
    assert best["segment"] in (0, 8, 12, 16, 20, 25, 28, 32) and best["ref_chunk"] > 0
    # the fastest candidate wins, except that per-read widths (segment 0)
    # stay unless a fixed width is at least 2% faster on the sample
    top = table[np.argmax(table["gcups"])]
    g0 = float(table[table["segment"] == 0]["gcups"][0])
    want = top["segment"] if top["gcups"] >= 1.02 * g0 else 0
    assert best["segment"] == want
    # the chunk is the plan's own rule for the whole workload at that width
    rule = ctx.plan(lns, precision="f16", segment=best["segment"])
    assert rule.info()["ref_chunk"] == best["ref_chunk"]
    rule.close()
    assert cache.read_text().split()[-2:] == [str(best["segment"]), str(best["ref_chunk"])]
    again, t2 = ctx.autotune(d_q, d_off, lns, sample_reads=16, sample_cols=8192, cache_path=cache)
    assert again == best and len(t2) == 0
    sc, am = ctx.score(reads, precision="f16", **best)
    sc0, am0 = ctx.score(reads, precision="f16", segment=32, ref_chunk=0)
    assert np.array_equal(sc, sc0) and np.array_equal(am, am0)


@pytest.mark.parametrize("segment", [8, 12, 16, 20, 25, 28, 32])
def test_f16_renorm_bit_exact(ctx, segment):
    """DTWAX_F16_RENORM equals the oracle's restatement of the renormalised
    form (drift at every lane boundary, the previous band's row minimum at
    lane 0, drift adapted band to band, offsets added back in binary32) bit
    for bit, scores and argmins, over reads of one to several bands."""
    rng = np.random.default_rng(600 + segment)
    ref = rng.standard_normal((2, 6000)).astype(np.float32)
    ctx.upload_reference(ref)
    BR = 64 * segment
    lens = [1, 100, BR - 1, BR, BR + 1, 2 * BR + 17, 3 * BR + 5, 5000]
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    n_t = min(3 * BR, 4500)  # a noisy cut of the reference, several bands long
    reads.append(ref[0, 1000:1000 + n_t] + rng.normal(0, 0.2, n_t).astype(np.float32))
    sc, am = ctx.score(reads, precision="f16_renorm", segment=segment)
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    osc, oam = O.batch_f16_renorm([O.f32_to_f16(r) for r in reads], sh, segment)
    assert np.array_equal(sc, osc), (sc, osc)
    assert np.array_equal(am, oam)


def test_f16_renorm_automatic_split_bit_exact(ctx):
    """The automatic renormalised plan: reads of up to 1536 rows in bands of
    64 x 8 rows, longer ones on the split schedule (their first 513..1536
    rows in bands of 64 x 8, the rest in bands of 64 x 16 at twice the
    adapted drift), equal to the oracle's restatement bit for bit."""
    rng = np.random.default_rng(650)
    ref = rng.standard_normal((2, 7000)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = [300, 1536, 1537, 1600, 2000, 2048, 2561, 3585, 5000]
    reads = [rng.standard_normal(n).astype(np.float32) for n in lens]
    reads.append(ref[1, 500:4500] + rng.normal(0, 0.2, 4000).astype(np.float32))
    sc, am = ctx.score(reads, precision="f16_renorm")
    sh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    osc, oam = O.batch_f16_renorm([O.f32_to_f16(r) for r in reads], sh, 0)
    assert np.array_equal(sc, osc), (sc, osc)
    assert np.array_equal(am, oam)


def test_autotune_with_packing_forced(ctx):
    """packing = 1 leaves the widths without a packed kernel (8, 12) out of
    the candidates instead of failing on them; the choice scores exactly."""
    rng = np.random.default_rng(78)
    ref = rng.standard_normal((2, 20000)).astype(np.float32)
    ctx.upload_reference(ref)
    lens = rng.integers(300, 5000, 40)
    reads = [rng.standard_normal(int(n)).astype(np.float32) for n in lens]
    flat, offs, lns = D.pack_reads(reads)
    d_q, d_off = D.DeviceArray.from_host(flat), D.DeviceArray.from_host(offs)
    best, table = ctx.autotune(d_q, d_off, lns, sample_reads=16, sample_cols=8192, packing=1)
    assert sorted(table["segment"].tolist()) == [0, 16, 20, 25, 28, 32]
    sc, am = ctx.score(reads, precision="f16", packing=1, **best)
    osc, oam = _oracle_f16(reads, ref)
    assert np.array_equal(sc, osc) and np.array_equal(am, oam)


def test_plan_and_score_argument_errors(ctx):
    """Bad options fail with a status and a message, not a fault."""
    rng = np.random.default_rng(4)
    ctx.upload_reference(rng.standard_normal((2, 1000)).astype(np.float32))
    for kw in (dict(segment=13), dict(ref_chunk=-1), dict(precision=7)):
        with pytest.raises(D.DtwaxError):
            ctx.plan(np.array([100, 200], np.int32), **kw)
    with pytest.raises(D.DtwaxError, match="at least one sample"):
        ctx.plan(np.array([100, 0], np.int32))
    q = [rng.standard_normal(50).astype(np.float32)]
    sc, am = ctx.score(q, precision="f16", segment=8)
    assert sc.shape == (1, 2) and np.all(np.isfinite(sc))


def test_waves_cap_is_exact(ctx):
    """Fewer resident waves (each pulling more work items) change nothing."""
    rng = np.random.default_rng(93)
    ref = rng.standard_normal((2, 3000)).astype(np.float32)
    ctx.upload_reference(ref)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (5, 700, 2100, 1300, 64, 4000)]
    base = ctx.score(reads, precision="f16", ref_chunk=500)
    for waves in (1, 3, 17):
        got = ctx.score(reads, precision="f16", ref_chunk=500, waves=waves)
        assert np.array_equal(base[0], got[0]) and np.array_equal(base[1], got[1])
    with pytest.raises(D.DtwaxError, match="waves"):
        ctx.plan(np.array([10], np.int32), waves=-1)
