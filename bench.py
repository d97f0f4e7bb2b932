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
reads, SIMD over columns) on a bounded sample of the same reads.
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--reads", type=int, default=4096)
    ap.add_argument("--precision", default="f16")
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the bounded cpu_baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="also run one plain device copy of 256 MiB (normaliser copy mode) "
                         "before timing, as the byte-count calibration for HBM counters")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sdtw_f16_kernel_48kb_4096x2000_traffic.json")


def measured_traffic():
    """HBM bytes per sDTW launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration)."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); "
              "the multi-GPU sharding is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)
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
    plan = ctx.plan(lens, precision=a.precision, segment=a.segment)
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
    best = res["score"].min(axis=1)
    thr = float(np.median(best))
    sep = bool(best[w.is_target].max() < best[~w.is_target].min())

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
                   "grid": pinfo["grid"]},
        "reads_classified_per_s": round(len(raw) / (step_ms * 1e-3), 1),
        "kernel_ms": round(kms, 3),
        "kernel_gcups": round(cell_rate / 1e9, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved_ops / 1e12, 2),
                     "peak": round(peak_ops / 1e12, 2), "unit": "TOP/s",
                     "frac": round(achieved_ops / peak_ops, 4),
                     "traffic": (measured_traffic() or {}).get("hbm_bytes_per_launch"),
                     "algorithmic_bytes": int(lens.sum()) * 4 + 2 * n_cols * 2 + len(raw) * 48,
                     "ops_per_cell": 3, "device": info},
        "checks": {"targets_separate": sep, "median_threshold": round(thr, 3)},
        "timing": {"wall_s": round(wall, 3), "gen_s": round(gen_s, 2)},
    }

    if not a.no_cpu:
        import oracle as O
        norm = d_norm.to_host()
        qs = [norm[o:o + n] for o, n in zip(offs, lens)]
        strands = w.ref_levels.astype(np.float32)
        probe = qs[:8]
        t1 = time.perf_counter()
        _, _, used = O.cpu_baseline_f32(probe, strands)
        tp = time.perf_counter() - t1
        n_sample = int(np.clip(len(probe) * a.cpu_seconds / max(tp, 1e-3), 8, len(qs)))
        sample = qs[:n_sample]
        t1 = time.perf_counter()
        sc, _, used = O.cpu_baseline_f32(sample, strands)
        tc = time.perf_counter() - t1
        ccells = sum(len(q) for q in sample) * n_cols * 2
        out["cpu_baseline"] = {"value": round(ccells / tc / 1e9, 2), "unit": "GCUPS", "cores": int(used),
                               "kind": "port",
                               "sample": f"{n_sample} of the {len(qs)} normalised reads, FP32, "
                                         f"{tc:.1f} s on {used} threads"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
