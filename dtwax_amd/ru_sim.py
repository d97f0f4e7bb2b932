"""Read Until serving simulation: a flow cell of channels streaming chunks
into the dtwax_ru_* engine.

The reference's own serving path is a prototype: ont_readuntil.py feeds the
last 16,000 bytes of every 2nd read into a FIFO and simulate_minion.py
writes 512 x 2048 random bytes every 0.5 s (src/ru_main.cu:172-326 scores
them statelessly and never answers).  This module is the load generator for
the stateful engine instead:

* n_channels channels sample at sample_rate Hz; reads start after an
  exponential gap (mean gap_s; with burst, every channel's first read
  starts at t = 0) and have exponential lengths (mean read_mean samples,
  at least one chunk); target_frac of them come from the
  reference (random strand and position), the rest from unrelated sequence;
* every chunk_len samples a read delivers one chunk, until it is decided;
  a NONTARGET read is unblocked (ends unblock_s after the decision), a
  TARGET read sequences to its natural end.

Three drivers:
* ``simulate`` -- the serving form, replayed event by event: whenever the
  server is idle it scores every chunk that has arrived (one per channel)
  with ReadUntil.process; simulated time advances by the measured wall time
  of the call (deterministic arrivals, real processing times);
* ``simulate_realtime`` -- the serving form in wall-clock time through the
  engine's own serving loop (submit/poll, overlapping batches on several
  streams); latencies are measured inside the engine;
* ``simulate_plain`` -- the plain batched form (the reference's way): a read
  is classified once, from scratch, after all max_chunks chunks have
  arrived, in batches of whatever reads are complete when the server is
  idle, with the batch scoring path.

Latency = decision available - chunk arrival (for the plain form: the
arrival of the read's last chunk).
"""
from __future__ import annotations

import heapq
import time
from dataclasses import dataclass

import numpy as np

from . import synth

CONTINUE, NONTARGET, TARGET = 0, 1, 2


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
    burst: bool = False  # every channel starts its first read at t = 0 (else after a gap)


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


def _best_cut(per_row, truth):
    """Midpoint between the classes if they separate, else the cut with the
    best balanced accuracy (target iff value < cut)."""
    t_hi = per_row[..., truth].max()
    n_lo = per_row[..., ~truth].min()
    if t_hi < n_lo:
        return float(0.5 * (t_hi + n_lo)), 1.0
    best, acc = float(n_lo), 0.0
    for c in np.unique(per_row):
        a = ((per_row[..., truth] < c).mean() + (per_row[..., ~truth] >= c).mean()) / 2
        if a > acc:
            best, acc = float(c), float(a)
    return best, acc


def calibrate_threshold_per_row(ctx, source, chunk_len, max_chunks, precision="f16", n_each=32,
                                norm=None):
    """Serving form: per-row threshold from n_each target and n_each
    non-target reads scored chunk by chunk by an engine with the threshold
    off.  Under the engine's rule a read is TARGET iff its per-row score
    stays below the threshold at every chunk, i.e. iff the largest per-row
    score over its chunks is below it, so the cut is made on that maximum.
    Returns (threshold_per_row, calibration accuracy)."""
    from . import NORM_MEDMAD, ReadUntil
    n = 2 * n_each
    truth = np.array([True] * n_each + [False] * n_each)
    sig = np.stack([source.signal(chunk_len * max_chunks, t) for t in truth])
    per_row = np.zeros((max_chunks, n))
    ru = ReadUntil(ctx, max_reads=n, chunk_len=chunk_len, max_chunks=max_chunks,
                   precision=precision, norm=NORM_MEDMAD if norm is None else norm)
    try:
        for k in range(max_chunks):
            dec = ru.process(np.arange(n), np.full(n, 1), sig[:, k * chunk_len:(k + 1) * chunk_len])
            per_row[k] = dec["score"].min(axis=1) / ((k + 1) * chunk_len)
    finally:
        ru.close()
    return _best_cut(per_row.max(axis=0), truth)


def _plain_scores(ctx, raws, precision, norm=None):
    """Plain batched form: statistics of each whole read, then the batch
    scoring path from scratch; min(fwd, rev) per read."""
    from . import NORM_MEDMAD
    normed, _ = ctx.normalize(raws, mode=NORM_MEDMAD if norm is None else norm)
    sc, _ = ctx.score(normed, precision=precision)
    return sc.min(axis=1)


def calibrate_plain_threshold(ctx, source, n_samples, precision="f16", n_each=32, norm=None):
    """Plain form: per-row threshold on whole n_samples reads."""
    truth = np.array([True] * n_each + [False] * n_each)
    raws = [source.signal(n_samples, t) for t in truth]
    per_row = _plain_scores(ctx, raws, precision, norm) / n_samples
    return _best_cut(per_row, truth)


class FlowCell:
    """Channels, reads and their event queue (start, chunk, end)."""

    def __init__(self, source, cfg: SimConfig):
        self.src, self.cfg = source, cfg
        self.rng = np.random.default_rng(cfg.seed)
        self.L, self.fs = cfg.chunk_len, cfg.sample_rate
        self.chunk_s = self.L / self.fs
        self.events, self.seq = [], 0
        self.reads = {}  # (channel, number) -> dict
        self.number = np.zeros(cfg.n_channels, np.int64)
        self.decided = []
        for ch in range(cfg.n_channels):
            self.push(0.0 if cfg.burst else float(self.rng.exponential(cfg.gap_s)), "start", ch)

    def push(self, t, kind, ch, payload=None):
        heapq.heappush(self.events, (t, self.seq, kind, ch, payload))
        self.seq += 1

    def start_read(self, ch, t0):
        cfg, L = self.cfg, self.L
        self.number[ch] += 1
        num = int(self.number[ch])
        n = max(L, int(self.rng.exponential(cfg.read_mean)))
        target = bool(self.rng.random() < cfg.target_frac)
        n_chunks = min(cfg.max_chunks, n // L)
        r = dict(t0=t0, n=n, target=target, end=t0 + n / self.fs, decided=None,
                 sig=self.src.signal(n_chunks * L, target), n_chunks=n_chunks, ended=False)
        self.reads[(ch, num)] = r
        for k in range(n_chunks):
            self.push(t0 + (k + 1) * self.chunk_s, "chunk", ch, (num, k))
        self.push(r["end"], "end", ch, num)

    def next_time(self):
        return self.events[0][0] if self.events else None

    def pop_due(self, now):
        """Yields due events after applying start/end bookkeeping:
        ("chunk", t, ch, num, k) for live reads, ("end", t, ch, num)."""
        while self.events and self.events[0][0] <= now:
            t, _, kind, ch, payload = heapq.heappop(self.events)
            if kind == "start":
                if t < self.cfg.duration_s:
                    self.start_read(ch, t)
            elif kind == "chunk":
                num, k = payload
                r = self.reads[(ch, num)]
                if r["decided"] is None and not r["ended"]:
                    yield ("chunk", t, ch, num, k)
            else:
                r = self.reads[(ch, payload)]
                if not r["ended"]:
                    r["ended"] = True
                    self.push(t + float(self.rng.exponential(self.cfg.gap_s)), "start", ch)
                    yield ("end", t, ch, payload)

    def decide(self, ch, num, action, now):
        r = self.reads[(ch, num)]
        if action not in (NONTARGET, TARGET) or r["decided"] is not None:
            return False
        r["decided"] = action
        self.decided.append(r)
        if action == NONTARGET:  # unblock: the read ends shortly after the decision
            new_end = now + self.cfg.unblock_s
            if new_end < r["end"] and not r["ended"]:
                r["end"] = new_end
                self.push(new_end, "end", ch, num)
        return True

    def done(self, now):
        nt = self.next_time()
        return nt is None or nt > self.cfg.duration_s + 5 * self.chunk_s

    def metrics(self, sim_s, lat, lat_dec, batch_sizes, busy):
        lat = np.asarray(lat) * 1e3
        lat_dec = np.asarray(lat_dec) * 1e3
        truth = np.array([r["target"] for r in self.decided], bool)
        call = np.array([r["decided"] == TARGET for r in self.decided], bool)

        def pct(a, q):
            return round(float(np.percentile(a, q)), 3) if a.size else None
        return dict(
            channels=self.cfg.n_channels, sim_seconds=round(sim_s, 2),
            reads_started=len(self.reads), reads_decided=len(self.decided),
            reads_decided_per_s=round(len(self.decided) / sim_s, 1),
            chunks_scored=int(lat.size), batches=len(batch_sizes),
            mean_batch=round(float(np.mean(batch_sizes)), 2) if batch_sizes else 0.0,
            max_batch=int(max(batch_sizes)) if batch_sizes else 0,
            latency_ms_p50=pct(lat, 50), latency_ms_p99=pct(lat, 99),
            latency_ms_max=round(float(lat.max()), 3) if lat.size else None,
            decision_latency_ms_p50=pct(lat_dec, 50), decision_latency_ms_p99=pct(lat_dec, 99),
            busy_frac=round(busy / sim_s, 4) if busy is not None else None,
            accuracy=round(float((truth == call).mean()), 4) if truth.size else None,
            recall_target=round(float(call[truth].mean()), 4) if truth.any() else None,
        )


def _one_per_channel(ready):
    ready.sort()
    take, keep, seen = [], [], set()
    for item in ready:
        if item[1] in seen:
            keep.append(item)
        else:
            seen.add(item[1])
            take.append(item)
    return take, keep


def simulate(ru, source, cfg: SimConfig):
    """Serving form, event by event (see the module docstring)."""
    fc, L = FlowCell(source, cfg), cfg.chunk_len
    now = 0.0
    ready = []  # (arrival, channel, number, k)
    lat, lat_dec, batch_sizes, busy = [], [], [], 0.0
    while True:
        for ev in fc.pop_due(now):
            if ev[0] == "chunk":
                ready.append(ev[1:])
            else:
                _, t, ch, num = ev
                ru.release(ch)
                ready = [it for it in ready if (it[1], it[2]) != (ch, num)]
        if not ready:
            if fc.done(now):
                break
            now = max(now, fc.next_time())
            continue
        take, ready = _one_per_channel(ready)
        sig = np.stack([fc.reads[(x[1], x[2])]["sig"][x[3] * L:(x[3] + 1) * L] for x in take])
        t0 = time.perf_counter()
        dec = ru.process([x[1] for x in take], [x[2] for x in take], sig)
        dt = time.perf_counter() - t0
        busy += dt
        now += dt
        batch_sizes.append(len(take))
        for x, d in zip(take, dec):
            lat.append(now - x[0])
            if fc.decide(x[1], x[2], int(d["action"]), now):
                lat_dec.append(now - x[0])
        ready = [it for it in ready if fc.reads[(it[1], it[2])]["decided"] is None]
    out = fc.metrics(max(now, cfg.duration_s), lat, lat_dec, batch_sizes, busy)
    st = ru.stats()
    out["gpu_cells_per_s_busy"] = round(st["cells"] / max(busy, 1e-9) / 1e12, 3)
    return out


def _take_decision(fc, d, now, lat, lat_dec):
    a = int(d["action"])
    if a > TARGET:  # NO_SLOT / STALE: answered without scoring
        return
    lat.append(d["latency_us"] * 1e-6)
    if fc.decide(int(d["channel"]), int(d["number"]), a, now):
        lat_dec.append(d["latency_us"] * 1e-6)


def simulate_realtime(ru, source, cfg: SimConfig):
    """Serving form in wall-clock time through the engine's serving loop."""
    fc, L = FlowCell(source, cfg), cfg.chunk_len
    ru.start()
    t_start = time.perf_counter()
    lat, lat_dec = [], []
    outstanding = 0
    try:
        while True:
            now = time.perf_counter() - t_start
            for d in ru.poll(4096, timeout_ms=0):
                outstanding -= 1
                _take_decision(fc, d, now, lat, lat_dec)
            for ev in fc.pop_due(now):
                if ev[0] == "chunk":
                    _, t, ch, num, k = ev
                    ru.submit(ch, num, fc.reads[(ch, num)]["sig"][k * L:(k + 1) * L])
                    outstanding += 1
                else:
                    ru.release(ev[2])
            if fc.done(now) and outstanding == 0:
                break
            nt = fc.next_time()
            wait = 0.002 if nt is None else min(0.002, max(0.0, nt - (time.perf_counter() - t_start)))
            if outstanding:
                got = ru.poll(4096, timeout_ms=int(wait * 1e3))
                now = time.perf_counter() - t_start
                for d in got:
                    outstanding -= 1
                    _take_decision(fc, d, now, lat, lat_dec)
            elif wait > 0:
                time.sleep(wait)
    finally:
        ru.stop()
    sim_s = max(time.perf_counter() - t_start, cfg.duration_s)
    st = ru.stats()
    out = fc.metrics(sim_s, lat, lat_dec, [], None)
    out["batches"] = st["batches"]
    out["mean_batch"] = round(st["chunks"] / max(st["batches"], 1), 2)
    out["gpu_kernel_s_sum"] = round(st["gpu_ms"] * 1e-3, 3)
    return out


def simulate_plain(ctx, source, cfg: SimConfig, threshold_per_row, precision="f16", norm=None):
    """Plain batched form, event by event (see the module docstring)."""
    fc, L = FlowCell(source, cfg), cfg.chunk_len
    need = cfg.max_chunks
    now = 0.0
    ready = []  # (arrival of the last chunk, channel, number)
    lat, lat_dec, batch_sizes, busy = [], [], [], 0.0
    while True:
        for ev in fc.pop_due(now):
            if ev[0] == "chunk":
                _, t, ch, num, k = ev
                if k == need - 1:
                    ready.append((t, ch, num))
            else:
                _, t, ch, num = ev
                ready = [it for it in ready if (it[1], it[2]) != (ch, num)]
        if not ready:
            if fc.done(now):
                break
            now = max(now, fc.next_time())
            continue
        take, ready = ready, []
        raws = [fc.reads[(x[1], x[2])]["sig"][:need * L] for x in take]
        t0 = time.perf_counter()
        best = _plain_scores(ctx, raws, precision, norm)
        dt = time.perf_counter() - t0
        busy += dt
        now += dt
        batch_sizes.append(len(take))
        for x, s in zip(take, best):
            a = NONTARGET if s >= threshold_per_row * need * L else TARGET
            lat.append(now - x[0])
            if fc.decide(x[1], x[2], a, now):
                lat_dec.append(now - x[0])
    return fc.metrics(max(now, cfg.duration_s), lat, lat_dec, batch_sizes, busy)
