#!/usr/bin/env python3
"""Benchmark: FP16 sDTW read classification on one MI355X.

Workload (the 48.5 kb FP16 drop-in configuration): a seeded random 48,500 bp
reference (48,495 k-mer levels per strand, both strands scored), 4096 reads
of 3000 raw samples trimmed by 1000 to 2000, 50% target / 50% non-target,
median/MAD normalised on the GPU.  Inputs are resident in HBM before the
timed region.  One step = normalise + score all 4096 reads (persistent sDTW
kernel, binary16 per-op arithmetic = the reference's half2 semantics) +
result finalise, on one stream.

Prints ONE JSON line.  `value` is whole-step GCUPS (cells = sum of read
length x counted columns x 2 strands); `roofline` prices the sDTW kernel's
own launch time (HIP events recorded on its stream) against the packed-VALU
peak N_CU x 64 lanes x 2 halves x f_clk, at 3 element ops per cell.
`cpu_baseline` times the FP32 CPU restatement (all host cores, OpenMP over
reads, SIMD over columns, column strips; built for the host's best ISA) on a
bounded sample of the same reads.

Beside the headline the default run also reports, each with its own
figures: the other arithmetic forms (`forms`); the reference's own launch
through dtwax_distances at its shipped geometry, with and without its seam
deviations (`reference_kernel`); one autotuned pass of the uneven workload
(16,384 reads of U[500, 8000] samples vs a 5 Mb reference; `uneven_5mb`,
with its own roofline and CPU baseline); and Read Until serving at 512
channels against the 5 Mb reference (`read_until`: a saturated batch, 60 s
of a served flow cell, the plain batched form, and a burst with every
channel starting at t = 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sDTW GCUPS + reads classified/s vs 48.5 kb ref; p99 Read Until decision latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 10 (c2), 1 (c4)")
    ap.add_argument("--warmup", type=int, default=None, help="default 2 (c2), 0 (c4)")
    ap.add_argument("--workload", default="c2", choices=("c2", "c4"),
                    help="c2: 48.5 kb x 4096 x 2000 FP16 (the headline); c4: 16,384 reads of "
                         "U[500, 8000] samples vs a 5 Mb reference, autotuned segment/chunk")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (c4)")
    ap.add_argument("--ref-chunk", type=int, default=0, help="c4 with --no-tune: reference chunk")
    ap.add_argument("--no-tune", action="store_true",
                    help="c4: use --segment (0 = per-read widths) and --ref-chunk instead of the autotuner")
    ap.add_argument("--reads", type=int, default=4096)
    ap.add_argument("--precision", default="f16")
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--packing", type=int, default=0, choices=(-1, 0, 1),
                    help="plan packing: 0 automatic, 1 always, -1 never")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the bounded cpu_baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ru", action="store_true", help="skip the Read Until serving leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the one-pass uneven 5 Mb leg")
    ap.add_argument("--ru-seconds", type=float, default=60.0,
                    help="simulated flow-cell seconds of the Read Until leg")
    ap.add_argument("--burst-seconds", type=float, default=10.0,
                    help="simulated seconds of the Read Until burst (every channel starts at t = 0)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="also run one plain device copy of 256 MiB (normaliser copy mode) "
                         "before timing, as the byte-count calibration for HBM counters")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sdtw_f16_kernel_48kb_4096x2000_traffic.json")


TRAFFIC_SOURCE = ("profiles/sdtw_f16_kernel_48kb_4096x2000_traffic.md: rocprofv3 --pmc FETCH_SIZE "
                  "and WRITE_SIZE of this launch, each scaled by a plain copy kernel's reading; "
                  "a committed measurement, not taken in this run")


def measured_traffic():
    """HBM bytes per sDTW launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration)."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def valu_peak_ops(ctx):
    """Packed-VALU element ops/s of the device: N_CU x 64 lanes x 2 halves x clock."""
    info = ctx.device_info()
    return info["n_cu"] * 64 * 2 * info["clock_khz"] * 1e3


def read_until_leg(ctx, seconds, norm="meanstd", burst_seconds=10.0):
    """Read Until serving on the 5 Mb workload: 512 channels, 1600-sample
    chunks (0.4 s at 4 kHz), decisions within 4 chunks, FP16, median/MAD of
    the first chunk ("meanstd": the reference's own mean/std on both the
    reference squiggle and the reads; "medmad": median/MAD on both).
    (1) a saturated batch: one first chunk on every channel
    at once; (2) a simulated flow cell driven in wall-clock time through the
    engine's serving loop (continuous batching on 2 streams), with the
    per-row threshold calibrated on 64 reads by the engine itself; (3) the
    same flow cell served by the plain batched form (each read classified
    once from scratch after its 4 chunks), threshold calibrated the same way."""
    import dtwax_amd as D
    from dtwax_amd import ru_sim, synth
    t0 = time.time()
    rng = np.random.default_rng(5)
    means, _ = synth.make_model(rng)
    seq, levels = synth.build_reference(rng, 5_000_000, means, ref_norm=norm)
    ctx.upload_reference(levels)
    mode = D.NORM_MEANSTD if norm == "meanstd" else D.NORM_MEDMAD
    src = ru_sim.ReadSource(seq, means, np.random.default_rng(6))
    thr, cal_acc = ru_sim.calibrate_threshold_per_row(ctx, src, 1600, 4, n_each=128, norm=mode)
    ru = D.ReadUntil(ctx, max_reads=512, chunk_len=1600, max_chunks=4, precision="f16",
                     threshold_per_row=thr, norm=mode)
    try:
        info = ru.info()
        sig = np.stack([src.signal(1600, i % 10 == 0) for i in range(512)])
        sat = []
        for rep in range(3):
            s0 = ru.stats()
            t1 = time.perf_counter()
            d_exact = ru.process(np.arange(512), np.full(512, rep + 1), sig)
            wall = time.perf_counter() - t1
            s1 = ru.stats()
            sat.append((wall, s1["gpu_ms"] - s0["gpu_ms"], s1["cells"] - s0["cells"]))
        for c in range(512):
            ru.release(c)
        wall, kms, cells = min(sat)
        print(f"[bench] read until: saturated batch {wall * 1e3:.1f} ms", file=sys.stderr, flush=True)
        # the opt-in fused form (one rounding of t*t + min) on the same chunks
        ruf = D.ReadUntil(ctx, max_reads=512, chunk_len=1600, max_chunks=4, precision="f16_fused",
                          threshold_per_row=thr, norm=mode)
        try:
            fsat = []
            for rep in range(3):
                s0 = ruf.stats()
                t1 = time.perf_counter()
                d_fused = ruf.process(np.arange(512), np.full(512, rep + 1), sig)
                fsat.append((time.perf_counter() - t1, ruf.stats()["gpu_ms"] - s0["gpu_ms"]))
            bf = None
            if burst_seconds > 0:
                for c in range(512):
                    ruf.release(c)
                bf = ru_sim.simulate_realtime(ruf, src, ru_sim.SimConfig(n_channels=512, duration_s=burst_seconds,
                                                                         seed=2, burst=True))
                print(f"[bench] read until: burst, fused form {bf['sim_seconds']} s", file=sys.stderr, flush=True)
        finally:
            ruf.close()
        fwall, fkms = min(fsat)
        fused_sat = {"wall_ms": round(fwall * 1e3, 2), "kernel_ms": round(fkms, 2),
                     "kernel_gcups": round(cells / (fkms * 1e-3) / 1e9, 1),
                     "actions_differing_from_exact": int(np.sum(d_fused["action"] != d_exact["action"]))}
        print(f"[bench] read until: saturated batch, fused form {fwall * 1e3:.1f} ms", file=sys.stderr, flush=True)
        cfg = ru_sim.SimConfig(n_channels=512, duration_s=seconds, seed=1)
        m = ru_sim.simulate_realtime(ru, src, cfg)
        print(f"[bench] read until: served {m['sim_seconds']} s", file=sys.stderr, flush=True)
        if burst_seconds > 0:
            for c in range(512):
                ru.release(c)
            bcfg = ru_sim.SimConfig(n_channels=512, duration_s=burst_seconds, seed=2, burst=True)
            b = ru_sim.simulate_realtime(ru, src, bcfg)
            print(f"[bench] read until: burst {b['sim_seconds']} s", file=sys.stderr, flush=True)
    finally:
        ru.close()
    pthr, pcal = ru_sim.calibrate_plain_threshold(ctx, src, 4 * 1600, n_each=128, norm=mode)
    plain = ru_sim.simulate_plain(ctx, src, cfg, pthr, norm=mode)
    plain.update({"threshold_per_row": pthr, "calibration_accuracy": pcal})
    print(f"[bench] read until: plain form {plain['sim_seconds']} s", file=sys.stderr, flush=True)
    m.update({
        "config": "5 Mb reference (2 x 4,999,995 columns), 512 channels, 1600-sample chunks "
                  f"(0.4 s at 4 kHz), <= 4 chunks per read, FP16, {norm} (first chunk), 10% target reads, "
                  "exponential read lengths (mean 20k samples) and gaps (mean 1 s)",
        "threshold_per_row": thr, "calibration_accuracy": cal_acc,
        "state_gb": round(info["state_bytes"] / 1e9, 2), "segment": info["segment"],
        "saturated_batch": {"chunks": 512, "wall_ms": round(wall * 1e3, 2),
                            "kernel_ms": round(kms, 2),
                            "kernel_gcups": round(cells / (kms * 1e-3) / 1e9, 1),
                            # 3 element ops per cell against packed VALU
                            "roofline_frac": round(3.0 * cells / (kms * 1e-3) / valu_peak_ops(ctx), 4),
                            "exact_form_ceiling_ms": round(4.0 * cells / valu_peak_ops(ctx) * 1e3, 1),
                            "chunks_per_s": round(512 / wall, 1),
                            "fused_form_opt_in": fused_sat},
        "plain_batched_form": plain,
        "serving_vs_plain_reads_per_s": round(m["reads_decided_per_s"] / max(plain["reads_decided_per_s"], 1e-9), 3),
        "leg_s": round(time.time() - t0, 1),  # the whole leg: set-up, calibration and every simulation
    })
    if burst_seconds > 0:
        m["burst"] = {"what": "every channel's first read starts at t = 0, so all 512 first chunks arrive "
                              "together (the saturated batch), then Poisson as above",
                      **{k: b[k] for k in ("sim_seconds", "reads_started", "reads_decided", "reads_decided_per_s",
                                           "latency_ms_p50", "latency_ms_p99", "latency_ms_max",
                                           "decision_latency_ms_p99", "mean_batch", "accuracy")}}
        m["burst"]["fused_form_opt_in"] = {k: bf[k] for k in ("reads_decided", "latency_ms_p50", "latency_ms_p99",
                                                              "latency_ms_max", "accuracy")}
    return m


def cpu_sample_rate(qs, strands, cpu_seconds):
    """FP32 CPU restatement (all host threads, OpenMP over reads) on a
    bounded sample of the normalised queries: the sample starts at one read
    per thread and grows until it takes about cpu_seconds."""
    import oracle as O
    n = min(len(qs), max(1, int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)))
    while True:
        sample = qs[:n]
        t1 = time.perf_counter()
        _, _, used = O.cpu_baseline_f32(sample, strands)
        tc = time.perf_counter() - t1
        if tc >= 0.5 * cpu_seconds or n >= len(qs):
            break
        n = min(len(qs), max(n + 1, int(n * cpu_seconds / max(tc, 1e-3))))
    ccells = sum(len(q) for q in sample) * strands.shape[1] * strands.shape[0]
    return ccells / tc, int(used), n, tc


def host_description():
    import oracle as O
    model, _ = O.cpu_flags()
    return {"cpu_model": model, "nproc": os.cpu_count(), "isa": O.baseline_lib()[1]}


def uneven_leg(a, steps, warmup, ctx=None):
    """16,384 reads, lengths uniform in [500, 8000] samples, 10% target, vs
    a 5 Mb random reference (2 x 4,999,995 columns), FP16, median/MAD.  The
    autotuner picks segment and reference chunk on a 512-read x 2^20-column
    sample; one step = one persistent launch over every (read, chunk) item.
    Returns the result object (the JSON line of --workload c4)."""
    import ctypes as C
    import dtwax_amd as D
    from dtwax_amd import synth
    t0 = time.time()
    w = synth.config_workload(seed=4, ref_bp=5_000_000, n_reads=16384, raw_len=None,
                              len_range=(500, 8000), target_frac=0.1)
    gen_s = time.time() - t0
    print(f"[bench] c4 workload generated in {gen_s:.1f} s", file=sys.stderr, flush=True)
    ctx = ctx or D.Context(0)
    info = ctx.device_info()
    ctx.upload_reference(w.ref_levels)
    n_cols = w.ref_levels.shape[1]
    flat, offs, lens = D.pack_reads(w.reads)
    d_raw, d_off, d_len = (D.DeviceArray.from_host(x) for x in (flat, offs, lens))
    d_norm = D.DeviceArray(flat.shape, np.float32)
    d_out = D.DeviceArray((len(lens),), D.RESULT_DTYPE)
    L = D.lib()
    st = C.c_void_p()
    D._check(L.dtwax_stream_create(C.byref(st)))
    stream = st.value
    D._check(L.dtwax_normalize(ctx.handle, d_raw.ptr, d_off.ptr, d_len.ptr, len(lens), D.NORM_MEDMAD,
                               0.0, d_norm.ptr, None, stream))
    D._check(L.dtwax_stream_sync(stream))
    t1 = time.time()
    if a.no_tune:
        best, table = dict(segment=a.segment, ref_chunk=a.ref_chunk), []
    else:
        best, table = ctx.autotune(d_norm, d_off, lens, sample_reads=512, sample_cols=1 << 20,
                                   cache_path=a.tune_cache, precision="f16")
    tune_s = time.time() - t1
    print(f"[bench] c4 autotune {best} in {tune_s:.1f} s", file=sys.stderr, flush=True)
    plan = ctx.plan(lens, precision="f16", packing=a.packing, **best)
    pinfo = plan.info()
    plan.set_timing(True)
    for _ in range(warmup):
        plan.run(d_norm, d_off, d_len, d_out, stream)
    D._check(L.dtwax_stream_sync(stream))
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    D._check(L.dtwax_event_create(C.byref(ev0)))
    D._check(L.dtwax_event_create(C.byref(ev1)))
    D._check(L.dtwax_device_sync())
    kms = []
    D._check(L.dtwax_event_record(ev0, stream))
    for i in range(steps):
        plan.run(d_norm, d_off, d_len, d_out, stream)
        kms.append(plan.kernel_ms())
        print(f"[bench] c4 step {i} kernel {kms[-1]:.1f} ms", file=sys.stderr, flush=True)
    D._check(L.dtwax_event_record(ev1, stream))
    D._check(L.dtwax_stream_sync(stream))
    ev_ms = C.c_float()
    D._check(L.dtwax_event_elapsed_ms(ev0, ev1, C.byref(ev_ms)))
    step_ms = ev_ms.value / steps
    res = d_out.to_host()
    bestsc = res["score"].min(axis=1) / lens  # per-row score: lengths differ
    order = np.sort(bestsc)
    accs = [((bestsc[w.is_target] < t).mean() + (bestsc[~w.is_target] >= t).mean()) / 2 for t in order]
    cells = pinfo["cells"]
    km = float(np.mean(kms))
    peak_ops = info["n_cu"] * 64 * 2 * info["clock_khz"] * 1e3
    achieved = 3.0 * cells / (km * 1e-3)
    out = {
        "metric": METRIC, "value": round(cells / (step_ms * 1e-3) / 1e9, 1), "unit": "GCUPS",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(step_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16",
        "data": "synthetic (seeded random 5 Mb reference, 6-mer model, simulated reads)",
        "config": {"workload": "16,384 reads x U[500, 8000] samples vs 5 Mb reference, both strands, "
                               "10% target, median/MAD, autotuned",
                   "reads": len(lens), "mean_query_len": round(float(lens.mean()), 1),
                   "ref_cols_per_strand": int(n_cols), "segment": pinfo["segment"],
                   "ref_chunk": pinfo["ref_chunk"], "work_items": pinfo["n_items"], "grid": pinfo["grid"],
                   "packed_bands": pinfo["packs"], "packed_reads": pinfo["packed_reads"]},
        "reads_classified_per_s": round(len(lens) / (step_ms * 1e-3), 1),
        "kernel_ms": round(km, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 2), "peak": round(peak_ops / 1e12, 2),
                     "unit": "TOP/s", "frac": round(achieved / peak_ops, 4), "traffic": None,
                     "ops_per_cell": 3, "computed_cells_over_counted": round(pinfo["computed"] / cells, 4),
                     "algorithmic_bytes": int(lens.sum()) * 4 + 2 * int(n_cols) * 2 + len(lens) * 48},
        "autotune": {"seconds": round(tune_s, 1), "choice": best,
                     "table": [{"segment": int(t["segment"]), "ref_chunk": int(t["ref_chunk"]),
                                "sample_gcups": round(float(t["gcups"]), 1)} for t in table]},
        "checks": {"balanced_accuracy_per_row_score": round(float(max(accs)), 4)},
        "timing": {"gen_s": round(gen_s, 1)},
    }
    if not a.no_cpu:
        norm = d_norm.to_host()
        rng = np.random.default_rng(0)
        pick = rng.permutation(len(lens))
        qs = [norm[offs[i]:offs[i] + lens[i]] for i in pick]
        rate, used, n_sample, tc = cpu_sample_rate(qs, w.ref_levels.astype(np.float32), a.cpu_seconds)
        out["cpu_baseline"] = {"value": round(rate / 1e9, 2), "unit": "GCUPS", "cores": used, "kind": "port",
                               "sample": f"{n_sample} random reads of the 16,384 (of "
                                         f"{int(sum(lens[pick[:n_sample]]))} samples) vs the whole "
                                         f"reference, FP32, {tc:.1f} s on {used} threads; the full "
                                         f"workload would take {cells / rate:.0f} s",
                               **host_description()}
    plan.close()
    for arr in (d_raw, d_off, d_len, d_norm, d_out):
        arr.free()
    return out


def main_c4(a):
    steps = 1 if a.steps is None else a.steps
    warmup = 0 if a.warmup is None else a.warmup
    print(json.dumps(uneven_leg(a, steps, warmup)))


def reference_kernel_leg(ctx, raw_reads, ref_levels, stream, peak_ops):
    """The reference's own launch at its shipped geometry through the
    FullDTW::distances mirror: 2048-sample queries (samples [952, 3000) of
    the headline's 3000-sample reads, mean/std with eps 1e-4 as
    normalizer.cu:86-100), the first 48,128 k-mers per strand as the tiled
    half2 reference (57 tiles = 47,424 columns scored), half2 dists.  Timed
    with HIP events around each call (untiling, keys, sDTW kernel, dists):
    compat_quirks 1 (the reference kernel's DP, seam deviations included,
    in the persistent kernel), 0 (the clean recurrence), 2 (the deviations
    from the tile-by-tile kernel); then the same through the FP32 build's
    mirror (dtwax_distances_f32)."""
    import ctypes as C
    import dtwax_amd as D
    L = D.lib()
    ref_len, tile, seg, warp = 48128, 832, 26, 32
    n_tiles = ref_len // tile
    normed, _ = ctx.normalize([r[952:3000] for r in raw_reads], mode=D.NORM_MEANSTD, eps=1e-4,
                              stream=stream)
    qs = np.stack(normed).astype(np.float32)
    n = qs.shape[0]

    def h2(x, y):
        a = np.asarray(x, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
        b = np.asarray(y, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
        return a | (b << 16)
    flat = h2(ref_levels[0, :ref_len], ref_levels[1, :ref_len])
    c = np.arange(n_tiles * tile)
    tiled = np.zeros(ref_len, np.uint32)
    tiled[(c // tile) * tile + (c % tile % seg) * warp + (c % tile) // seg] = flat[c]
    words = np.concatenate([h2(qs.ravel(), qs.ravel()), np.zeros(32, np.uint32)])
    d_ref, d_q = D.DeviceArray.from_host(tiled), D.DeviceArray.from_host(words)
    d_out = D.DeviceArray((n,), np.uint32)
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    D._check(L.dtwax_event_create(C.byref(ev0)))
    D._check(L.dtwax_event_create(C.byref(ev1)))
    cells = n * 2048 * n_tiles * tile * 2
    out, scores = {}, {}
    for name, quirks in (("reference_dp", 1), ("clean", 0), ("reference_dp_tiles", 2)):
        ms = []
        for _ in range(2):  # the first call warms up
            D._check(L.dtwax_event_record(ev0, stream))
            ctx.distances(d_ref, d_q, d_out, n, compat_quirks=quirks, stream=stream)
            D._check(L.dtwax_event_record(ev1, stream))
            D._check(L.dtwax_stream_sync(stream))
            v = C.c_float()
            D._check(L.dtwax_event_elapsed_ms(ev0, ev1, C.byref(v)))
            ms.append(v.value)
        scores[quirks] = d_out.to_host().view(np.float16).reshape(n, 2).astype(np.float32)
        out[name] = {"compat_quirks": quirks, "call_ms": round(ms[-1], 3),
                     "gcups": round(cells / (ms[-1] * 1e-3) / 1e9, 1),
                     "roofline_frac": round(3.0 * cells / (ms[-1] * 1e-3) / peak_ops, 4)}
    # the FP32 build (FullDTW::distances<float>): the same queries against
    # fwd | rev as 96,256 floats, 115 tiles = 95,680 columns of one sequence
    cat = np.concatenate([ref_levels[0, :ref_len], ref_levels[1, :ref_len]]).astype(np.float32)
    n32_tiles = 2 * ref_len // tile
    c32 = np.arange(n32_tiles * tile)
    tiled32 = np.zeros(2 * ref_len, np.float32)
    tiled32[(c32 // tile) * tile + (c32 % tile % seg) * warp + (c32 % tile) // seg] = cat[c32]
    d_ref32 = D.DeviceArray.from_host(tiled32)
    d_q32 = D.DeviceArray.from_host(np.concatenate([qs.ravel(), np.zeros(32, np.float32)]))
    d_out32 = D.DeviceArray((n,), np.float32)
    cells32 = n * 2048 * n32_tiles * tile
    fp32, scores32 = {}, {}
    for name, quirks in (("reference_dp", 1), ("clean", 0), ("reference_dp_tiles", 2)):
        ms = []
        for _ in range(2):
            D._check(L.dtwax_event_record(ev0, stream))
            ctx.distances_f32(d_ref32, d_q32, d_out32, n, compat_quirks=quirks, stream=stream)
            D._check(L.dtwax_event_record(ev1, stream))
            D._check(L.dtwax_stream_sync(stream))
            v = C.c_float()
            D._check(L.dtwax_event_elapsed_ms(ev0, ev1, C.byref(v)))
            ms.append(v.value)
        scores32[quirks] = d_out32.to_host()
        fp32[name] = {"compat_quirks": quirks, "call_ms": round(ms[-1], 3),
                      "gcups": round(cells32 / (ms[-1] * 1e-3) / 1e9, 1),
                      "roofline_frac": round(3.0 * cells32 / (ms[-1] * 1e-3) / peak_ops, 4)}
    d32 = scores32[1] - scores32[0]
    fp32.update({
        "workload": f"the same {n} queries vs fwd | rev as 96,256 floats, 95,680 columns of one sequence",
        "tiles_equal_persistent": bool(np.array_equal(scores32[1], scores32[2])),
        "clean_vs_reference_dp": {"changed_scores": int(np.count_nonzero(d32)),
                                  "max_abs_diff": float(np.abs(d32).max()),
                                  "mean_abs_diff": float(f"{np.abs(d32).mean():.4g}")},
    })
    out["fp32_build"] = fp32
    L.dtwax_event_destroy(ev0)
    L.dtwax_event_destroy(ev1)
    d = scores[1] - scores[0]
    best_r, best_c = scores[1].min(axis=1), scores[0].min(axis=1)
    cuts = np.unique(best_r)
    out.update({
        "workload": f"{n} reads x 2048 samples (mean/std, eps 1e-4) vs the tiled half2 reference of "
                    "48,128 k-mers per strand, 47,424 columns scored, both strands",
        "tiles_equal_persistent": bool(np.array_equal(scores[1], scores[2])),
        "clean_vs_reference_dp": {"changed_strand_scores": int(np.count_nonzero(d)),
                                  "max_abs_diff": float(np.abs(d).max()),
                                  "mean_abs_diff": float(f"{np.abs(d).mean():.4g}"),
                                  "max_decision_flips_over_all_thresholds":
                                      int(max(np.sum((best_r < t) != (best_c < t)) for t in cuts))},
    })
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): the path's sharding over "
              "several GPUs (reads or reference chunks per GPU, one min-reduction) is not built",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "c4":
        return main_c4(a)
    a.steps = 10 if a.steps is None else a.steps
    a.warmup = 2 if a.warmup is None else a.warmup
    import dtwax_amd as D
    from dtwax_amd import synth

    t0 = time.time()
    w = synth.config_workload(seed=2, ref_bp=48_500, n_reads=a.reads, raw_len=3000, target_frac=0.5)
    raw = synth.trim_reads(w.reads, 1000)
    gen_s = time.time() - t0

    ctx = D.Context(0)
    info = ctx.device_info()
    ctx.upload_reference(w.ref_levels)
    n_cols = w.ref_levels.shape[1]
    flat, offs, lens = D.pack_reads(raw)
    d_raw = D.DeviceArray.from_host(flat)
    d_off = D.DeviceArray.from_host(offs)
    d_len = D.DeviceArray.from_host(lens)
    d_norm = D.DeviceArray(flat.shape, np.float32)
    d_out = D.DeviceArray((len(raw),), D.RESULT_DTYPE)
    plan = ctx.plan(lens, precision=a.precision, segment=a.segment, packing=a.packing)
    pinfo = plan.info()
    plan.set_timing(True)
    L = D.lib()

    import ctypes as C
    st = C.c_void_p()
    D._check(L.dtwax_stream_create(C.byref(st)))
    stream = st.value

    def step():
        D._check(L.dtwax_normalize(ctx.handle, d_raw.ptr, d_off.ptr, d_len.ptr, len(raw),
                                   D.NORM_MEDMAD, 0.0, d_norm.ptr, None, stream))
        plan.run(d_norm, d_off, d_len, d_out, stream)

    if a.traffic_probe:
        n_copy = 64 << 20  # floats: 256 MiB read + 256 MiB written
        d_a = D.DeviceArray((n_copy,), np.float32)
        d_b = D.DeviceArray((n_copy,), np.float32)
        d_o = D.DeviceArray.from_host(np.zeros(1, np.int64))
        d_l = D.DeviceArray.from_host(np.array([n_copy], np.int32))
        D._check(L.dtwax_normalize(ctx.handle, d_a.ptr, d_o.ptr, d_l.ptr, 1, D.NORM_NONE, 0.0,
                                   d_b.ptr, None, stream))
        D._check(L.dtwax_stream_sync(stream))

    for _ in range(a.warmup):
        step()
    D._check(L.dtwax_stream_sync(stream))
    kernel_ms = []
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    D._check(L.dtwax_event_create(C.byref(ev0)))
    D._check(L.dtwax_event_create(C.byref(ev1)))
    D._check(L.dtwax_device_sync())
    t_start = time.perf_counter()
    D._check(L.dtwax_event_record(ev0, stream))
    for _ in range(a.steps):
        step()
        kernel_ms.append(plan.kernel_ms())  # waits for this step's sDTW kernel
    D._check(L.dtwax_event_record(ev1, stream))
    D._check(L.dtwax_stream_sync(stream))
    D._check(L.dtwax_device_sync())
    wall = time.perf_counter() - t_start
    ev_ms = C.c_float()
    D._check(L.dtwax_event_elapsed_ms(ev0, ev1, C.byref(ev_ms)))
    step_ms = ev_ms.value / a.steps

    res = d_out.to_host()
    # the same step from host buffers to host results: the raw reads copied in
    # from pageable host memory (synchronous), normalised and scored, the
    # results copied back; best of 3 (wall clock)
    pcie = []
    for _ in range(3):
        D._check(L.dtwax_device_sync())
        t1 = time.perf_counter()
        D._check(L.dtwax_memcpy_h2d(d_raw.ptr, flat.ctypes.data, flat.nbytes, stream))
        step()
        host_res = d_out.to_host(stream)  # on the step's stream, after its kernels
        pcie.append(time.perf_counter() - t1)
    assert np.array_equal(host_res["score"], res["score"])
    pcie_ms = min(pcie) * 1e3
    best = res["score"].min(axis=1)
    # accuracy of the best single threshold (target iff score < t)
    order = np.sort(best)
    cuts = np.concatenate([order, [np.inf]])
    accs = [((best[w.is_target] < t).mean() + (best[~w.is_target] >= t).mean()) / 2 for t in cuts]
    thr = float(cuts[int(np.argmax(accs))])
    acc = float(max(accs))

    cells = pinfo["cells"]
    gcups = cells / (step_ms * 1e-3) / 1e9
    kms = float(np.mean(kernel_ms))
    cell_rate = cells / (kms * 1e-3)
    peak_ops = info["n_cu"] * 64 * 2 * info["clock_khz"] * 1e3
    achieved_ops = 3.0 * cell_rate

    out = {
        "metric": METRIC,
        "value": round(gcups, 1),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.precision,
        "data": "synthetic (seeded random 48.5 kb reference, 6-mer model, simulated reads)",
        "config": {"workload": "48.5 kb reference x 4096 reads x 2000 samples (3000 raw, 1000 trimmed), "
                               "both strands, 50% target, median/MAD",
                   "reads": len(raw), "query_len": int(lens[0]), "ref_cols_per_strand": int(n_cols),
                   "segment": pinfo["segment"], "work_items": pinfo["n_items"],
                   "grid": pinfo["grid"], "packed_bands": pinfo["packs"], "packed_reads": pinfo["packed_reads"]},
        "reads_classified_per_s": round(len(raw) / (step_ms * 1e-3), 1),
        "pcie_inclusive": {"what": "host buffers to host results: raw reads copied from pageable host memory, "
                                   "normalised, scored, results copied back (not the value)",
                           "ms": round(pcie_ms, 3), "gcups": round(cells / (pcie_ms * 1e-3) / 1e9, 1),
                           "h2d_bytes": int(flat.nbytes),
                           "reads_classified_per_s": round(len(raw) / (pcie_ms * 1e-3), 1)},
        # the reference's own throughput figure (hpc_helpers.hpp:93-100): query
        # samples per microsecond of the sDTW launch / 10 (~10 samples per base)
        "reference_mbps": round(float(lens.sum()) / (kms * 1e3) / 10, 1),
        "kernel_ms": round(kms, 3),
        "kernel_gcups": round(cell_rate / 1e9, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved_ops / 1e12, 2),
                     "peak": round(peak_ops / 1e12, 2), "unit": "TOP/s",
                     "frac": round(achieved_ops / peak_ops, 4),
                     "traffic": (measured_traffic() or {}).get("hbm_bytes_per_launch"),
                     "traffic_source": TRAFFIC_SOURCE,
                     "algorithmic_bytes": int(lens.sum()) * 4 + 2 * n_cols * 2 + len(raw) * 48,
                     "ops_per_cell": 3, "device": info},
        "checks": {"balanced_accuracy": round(acc, 4), "threshold": round(thr, 3)},
        "timing": {"wall_s": round(wall, 3), "gen_s": round(gen_s, 2)},
    }

    # the other arithmetic forms on the same normalised reads (kernel time of
    # one timed launch after one warm-up; FP16_FUSED rounds t*t + min once,
    # FP16_RENORM subtracts drift, FP32 is the reference's FP32 build;
    # "f16_moves3" is the reference's build without -DNO_REF_DEL, whose cell
    # also takes the left neighbour, over whole strands)
    forms = {}
    for name, prec, moves in (("f16_fused", "f16_fused", 2), ("f16_renorm", "f16_renorm", 2),
                              ("f32", "f32", 2), ("f16_moves3", "f16", 3)):
        if prec == a.precision and moves == 2:
            continue
        fp = ctx.plan(lens, precision=prec, moves=moves)
        fp.set_timing(True)
        fms = []
        for _ in range(2):
            fp.run(d_norm, d_off, d_len, d_out, stream)
            fms.append(fp.kernel_ms())
        fi = fp.info()
        # against the headline form on the same reads: score difference and
        # decisions at the headline's threshold
        fs = d_out.to_host()["score"]
        rel = np.abs(fs - res["score"]) / np.maximum(np.abs(res["score"]), 1e-30)
        flips = int(np.sum((fs.min(axis=1) < thr) != (best < thr)))
        forms[name] = {"kernel_ms": round(fms[-1], 3), "segment": fi["segment"],
                       "kernel_gcups": round(fi["cells"] / (fms[-1] * 1e-3) / 1e9, 1),
                       "roofline_frac": round(3.0 * fi["cells"] / (fms[-1] * 1e-3) / peak_ops, 4),
                       "vs_headline": {"max_rel_score_diff": float(f"{rel.max():.3g}"),
                                       "mean_rel_score_diff": float(f"{rel.mean():.3g}"),
                                       "decision_flips": flips}}
        fp.close()
    out["forms"] = forms
    out["reference_kernel"] = reference_kernel_leg(ctx, w.reads, w.ref_levels, stream, peak_ops)
    print("[bench] reference launch done", file=sys.stderr, flush=True)

    if not a.no_cpu:
        norm = d_norm.to_host()
        qs = [norm[o:o + n] for o, n in zip(offs, lens)]
        rate, used, n_sample, tc = cpu_sample_rate(qs, w.ref_levels.astype(np.float32), a.cpu_seconds)
        out["cpu_baseline"] = {"value": round(rate / 1e9, 2), "unit": "GCUPS", "cores": used,
                               "kind": "port",
                               "sample": f"{n_sample} of the {len(qs)} normalised reads, FP32, "
                                         f"{tc:.1f} s on {used} threads",
                               **host_description()}
    plan.close()
    for arr in (d_raw, d_off, d_len, d_norm, d_out):
        arr.free()
    if not a.no_c4:
        print("[bench] uneven 5 Mb leg", file=sys.stderr, flush=True)
        c4 = uneven_leg(a, 1, 0, ctx)
        out["uneven_5mb"] = {k: c4[k] for k in ("value", "unit", "ms_per_step", "kernel_ms", "reads_classified_per_s",
                                                "config", "roofline", "autotune", "checks")}
        if "cpu_baseline" in c4:
            out["uneven_5mb"]["cpu_baseline"] = c4["cpu_baseline"]
    if not a.no_ru:
        out["read_until"] = read_until_leg(ctx, a.ru_seconds, burst_seconds=a.burst_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
