"""Read Until serving simulation: a flow cell of channels streaming chunks
into the dtwax_ru_* engine, replayed event by event.

The reference's own serving path is a prototype: ont_readuntil.py feeds the
last 16,000 bytes of every 2nd read into a FIFO and simulate_minion.py
writes 512 x 2048 random bytes every 0.5 s (src/ru_main.cu:172-326 scores
them statelessly and never answers).  This module is the load generator for
the stateful engine instead:

* n_channels channels sample at sample_rate Hz; reads start after an
  exponential gap (mean gap_s) and have exponential lengths (mean
  read_mean samples, at least one chunk); target_frac of them come from the
  reference (random strand and position), the rest from unrelated sequence;
* every chunk_len samples a read delivers one chunk, until the engine
  decides it; a NONTARGET read is unblocked (ends unblock_s after the
  decision), a TARGET read sequences to its natural end;
* the server takes every chunk that has arrived whenever it is idle (one
  chunk per channel per batch) and calls ReadUntil.process; simulated time
  advances by the measured wall time of that call, so latencies are real
  GPU + host processing times with simulated arrivals.

Latency = decision available - chunk arrival.
"""
from __future__ import annotations

import heapq
import time
from dataclasses import dataclass

import numpy as np

from . import synth


@dataclass
class SimConfig:
    n_channels: int = 512
    duration_s: float = 20.0
    sample_rate: float = 4000.0
    chunk_len: int = 1600
    max_chunks: int = 4
    read_mean: float = 20000.0
    gap_s: float = 1.0
    unblock_s: float = 0.1
    target_frac: float = 0.1
    seed: int = 0


class ReadSource:
    """Raw signal of the first max_chunks chunks of simulated reads."""

    def __init__(self, seq, model_means, rng, k=6):
        self.means = model_means
        self.rng = rng
        self.k = k
        self.codes = (synth.kmer_codes(seq, k), synth.kmer_codes(synth.revcomp(seq), k))

    def signal(self, n_samples, target):
        span = int(n_samples / 9.0 * 1.5) + 64
        if target:
            codes = self.codes[int(self.rng.integers(0, 2))]
            start = int(self.rng.integers(0, max(1, codes.size - span)))
            lv = self.means[codes[start:start + span]]
        else:
            other = synth.random_seq(self.rng, span + self.k)
            lv = self.means[synth.kmer_codes(other, self.k)]
        return synth.simulate_signal(self.rng, lv, n_samples)


def calibrate_threshold_per_row(ctx, source, chunk_len, max_chunks, precision="f16", n_each=32):
    """Per-row threshold from n_each target and n_each non-target reads
    scored chunk by chunk by an engine with the threshold off: the midpoint
    between the largest target and the smallest non-target score per row
    over every chunk count, or the accuracy-maximising cut if they overlap.
    Returns (threshold_per_row, calibration accuracy)."""
    from . import ReadUntil
    n = 2 * n_each
    truth = np.array([True] * n_each + [False] * n_each)
    sig = np.stack([source.signal(chunk_len * max_chunks, t) for t in truth])
    per_row = np.zeros((max_chunks, n))
    ru = ReadUntil(ctx, max_reads=n, chunk_len=chunk_len, max_chunks=max_chunks,
                   precision=precision)
    try:
        for k in range(max_chunks):
            dec = ru.process(np.arange(n), np.full(n, 1), sig[:, k * chunk_len:(k + 1) * chunk_len])
            per_row[k] = dec["score"].min(axis=1) / ((k + 1) * chunk_len)
    finally:
        ru.close()
    t_hi = per_row[:, truth].max()
    n_lo = per_row[:, ~truth].min()
    if t_hi < n_lo:
        return float(0.5 * (t_hi + n_lo)), 1.0
    cand = np.unique(per_row)
    best, acc = float(cand[0]), 0.0
    for c in cand:
        a = ((per_row[:, truth] < c).mean() + (per_row[:, ~truth] >= c).mean()) / 2
        if a > acc:
            best, acc = float(c), a
    return best, float(acc)


def simulate(ru, source, cfg: SimConfig):
    """Runs the flow cell for cfg.duration_s simulated seconds; returns metrics."""
    rng = np.random.default_rng(cfg.seed)
    L, fs = cfg.chunk_len, cfg.sample_rate
    chunk_s = L / fs
    events = []  # (time, seq, kind, channel, payload)
    seq = 0

    def push(t, kind, ch, payload=None):
        nonlocal seq
        heapq.heappush(events, (t, seq, kind, ch, payload))
        seq += 1

    reads = {}  # (channel, number) -> dict
    number = np.zeros(cfg.n_channels, np.int64)

    def start_read(ch, t0):
        number[ch] += 1
        n = max(L, int(rng.exponential(cfg.read_mean)))
        target = bool(rng.random() < cfg.target_frac)
        n_chunks = min(cfg.max_chunks, n // L)
        r = dict(t0=t0, n=n, target=target, end=t0 + n / fs, decided=None, t_dec=None,
                 sig=source.signal(n_chunks * L, target), n_chunks=n_chunks)
        reads[(ch, int(number[ch]))] = r
        for k in range(n_chunks):
            push(t0 + (k + 1) * chunk_s, "chunk", ch, (int(number[ch]), k))
        push(r["end"], "end", ch, int(number[ch]))

    for ch in range(cfg.n_channels):
        push(float(rng.exponential(cfg.gap_s)), "start", ch)

    now = 0.0
    ready = []  # (arrival, channel, number, k)
    lat, lat_dec, batch_sizes, busy = [], [], [], 0.0
    decided = []
    while True:
        while events and events[0][0] <= now:
            t, _, kind, ch, payload = heapq.heappop(events)
            if kind == "start":
                if t < cfg.duration_s:
                    start_read(ch, t)
            elif kind == "chunk":
                num, k = payload
                r = reads[(ch, num)]
                if r["decided"] is None and not r.get("ended"):
                    ready.append((t, ch, num, k))
            elif kind == "end":
                r = reads[(ch, payload)]
                if not r.get("ended"):
                    r["ended"] = True
                    ru.release(ch)
                    ready = [it for it in ready if (it[1], it[2]) != (ch, payload)]
                    push(t + float(rng.exponential(cfg.gap_s)), "start", ch)
        if not ready:
            if not events or events[0][0] > cfg.duration_s + 5 * chunk_s:
                break
            now = max(now, events[0][0])
            continue
        # one chunk per channel, earliest first; the rest wait
        ready.sort()
        take, keep, seen = [], [], set()
        for item in ready:
            if item[1] in seen:
                keep.append(item)
            else:
                seen.add(item[1])
                take.append(item)
        ready = keep
        chans = np.array([x[1] for x in take], np.int32)
        nums = np.array([x[2] for x in take], np.int32)
        sig = np.stack([reads[(x[1], x[2])]["sig"][x[3] * L:(x[3] + 1) * L] for x in take])
        t0 = time.perf_counter()
        dec = ru.process(chans, nums, sig)
        dt = time.perf_counter() - t0
        busy += dt
        now += dt
        batch_sizes.append(len(take))
        for x, d in zip(take, dec):
            lat.append(now - x[0])
            r = reads[(x[1], x[2])]
            a = int(d["action"])
            if a in (1, 2) and r["decided"] is None:
                r["decided"] = a
                r["t_dec"] = now
                lat_dec.append(now - x[0])
                decided.append(r)
                if a == 1:  # unblock: the read ends shortly after the decision
                    new_end = now + cfg.unblock_s
                    if new_end < r["end"]:
                        r["end"] = new_end
                        push(new_end, "end", x[1], x[2])
            # ready chunks of a decided read are dropped
        ready = [it for it in ready if reads[(it[1], it[2])]["decided"] is None]
    lat = np.array(lat) * 1e3
    lat_dec = np.array(lat_dec) * 1e3
    truth = np.array([r["target"] for r in decided])
    call = np.array([r["decided"] == 2 for r in decided])
    sim_s = max(now, cfg.duration_s)
    st = ru.stats()
    out = dict(
        channels=cfg.n_channels, sim_seconds=round(sim_s, 2),
        reads_started=len(reads), reads_decided=len(decided),
        reads_decided_per_s=round(len(decided) / sim_s, 1),
        chunks_scored=int(len(lat)), batches=len(batch_sizes),
        mean_batch=round(float(np.mean(batch_sizes)) if batch_sizes else 0.0, 2),
        max_batch=int(max(batch_sizes)) if batch_sizes else 0,
        latency_ms_p50=round(float(np.percentile(lat, 50)), 3) if lat.size else None,
        latency_ms_p99=round(float(np.percentile(lat, 99)), 3) if lat.size else None,
        latency_ms_max=round(float(lat.max()), 3) if lat.size else None,
        decision_latency_ms_p99=round(float(np.percentile(lat_dec, 99)), 3) if lat_dec.size else None,
        busy_frac=round(busy / sim_s, 4),
        gpu_cells_per_s_busy=round(st["cells"] / max(st["gpu_ms"] * 1e-3, 1e-9) / 1e12, 3),
        accuracy=round(float((truth == call).mean()), 4) if truth.size else None,
        recall_target=round(float(call[truth].mean()), 4) if truth.any() else None,
    )
    return out
