"""A simulated sequencer behind MinKNOW's data service, for dry runs of the
Read Until client (`dtwax_amd.minknow`) without a device.

The reference's stand-in, src/ReadUntil/simulate_minion.py, writes random
bytes into the classifier's FIFO.  This one serves a flow cell over the same
gRPC methods MinKNOW exposes (get_live_reads, get_data_types) and enacts the
actions it receives:

* channels first..last sample at sample_rate Hz (time runs `speed` times
  faster than the wall clock); each channel alternates an exponential gap
  (mean gap_s) and a read of exponential length (mean read_mean samples);
  target_frac of the reads come from the reference (random strand and
  position), the rest from unrelated random sequence; the signal is
  generated as it is sequenced (k-mer levels, geometric dwell, Gaussian
  noise), so long reads cost nothing until they are sent;
* every interval_s of wall time one GetLiveReadsResponse carries, per
  channel with a live read whose data is wanted, the samples acquired since
  the last message (calibrated float32), plus the responses to the actions
  received meanwhile;
* unblock ends the read at once; stop_further_data stops its data (the read
  sequences to its natural end); an action for a read that has ended, or
  for another read number, fails with FAILED_READ_FINISHED.

`summary()` reports what sequencing yielded per class (reads, unblocked,
samples sequenced), from which enrichment follows.

Playback: with `playback` (a POD5 path or `Pod5File`), each new read is the
next read of the file, in table order and cycling, sent as calibrated
picoamperes ((raw + calibration_offset) x calibration_scale) under its own
read id and at its own length; `playback_truth` ({read_id: is_target})
labels them for `summary()` (unlisted reads count as non-target).
"""
from __future__ import annotations

import threading
import time
import uuid

import numpy as np

from . import minknow as mk
from . import synth


class _Read:
    __slots__ = ("number", "id", "start", "length", "target", "codes", "pos", "pending",
                 "sent", "stopped", "end", "unblocked")

    def __init__(self, number, start, length, target, codes):
        self.number, self.start, self.length, self.target = number, start, length, target
        self.id = str(uuid.uuid4())
        self.codes, self.pos = codes, 0
        self.pending = np.zeros(0, np.float32)  # generated, not yet sent
        self.sent = start
        self.stopped = False
        self.end = start + length
        self.unblocked = False


class SimulatedSequencer:
    """DataService servicer (register with minknow.add_data_service)."""

    def __init__(self, seq, model_means, n_channels=512, sample_rate=4000.0, read_mean=20000.0,
                 gap_s=1.0, target_frac=0.1, speed=1.0, interval_s=0.1, duration_s=None,
                 min_read=1600, seed=0, k=6, playback=None, playback_truth=None):
        self.means = np.asarray(model_means, np.float32)
        self.k = k
        self.strands = (synth.kmer_codes(seq, k), synth.kmer_codes(synth.revcomp(seq), k))
        self.n_channels = n_channels
        self.fs = sample_rate
        self.read_mean, self.gap_s, self.target_frac = read_mean, gap_s, target_frac
        self.speed, self.interval_s, self.duration_s = speed, interval_s, duration_s
        self.min_read = min_read
        self.rng = np.random.default_rng(seed)
        self.lock = threading.Lock()
        self.finished = []  # ended reads
        self.reads = {}  # (channel, number) -> every read started
        self.n_streams = 0
        self._pb = None
        if playback is not None:
            from . import Pod5File
            self._pb = playback if isinstance(playback, Pod5File) else Pod5File(playback)
            self._pb_info = self._pb.reads()
            self._pb_order = [i for i in range(self._pb.n_reads) if self._pb_info["n_samples"][i] > 0]
            if not self._pb_order:
                raise ValueError("the POD5 file holds no read with samples")
            self._pb_next = 0
            self._pb_truth = dict(playback_truth or {})

    # ---- flow cell ---------------------------------------------------------
    def _playback_read(self, number, start):
        i = self._pb_order[self._pb_next % len(self._pb_order)]
        self._pb_next += 1
        inf = self._pb_info[i]
        x = self._pb.signal(reads=[i])[0].astype(np.float32)
        off, sc = inf["calibration_offset"], inf["calibration_scale"]
        if np.isfinite(off) and np.isfinite(sc):
            x = (x + np.float32(off)) * np.float32(sc)
        rid = inf["read_id"].decode()
        r = _Read(number, start, x.size, bool(self._pb_truth.get(rid, False)), None)
        r.id = rid
        r.pending = x
        return r

    def _new_read(self, number, start):
        if self._pb is not None:
            return self._playback_read(number, start)
        length = max(self.min_read, int(self.rng.exponential(self.read_mean)))
        target = bool(self.rng.random() < self.target_frac)
        if target:
            codes = self.strands[int(self.rng.integers(0, 2))]
            span = int(length / 9.0 * 1.5) + 64
            p0 = int(self.rng.integers(0, max(1, codes.size - span)))
            codes = codes[p0:]
        else:
            codes = None  # unrelated sequence, drawn as it is needed
        return _Read(number, start, length, target, codes)

    def _signal(self, r, n):
        """The read's next n samples (continuing its k-mer walk, or the
        played-back signal)."""
        while self._pb is None and r.pending.size < n:
            block = max(n - r.pending.size, 4000)
            span = int(block / 9.0 * 1.5) + 64
            if r.codes is None:
                lv = self.means[synth.kmer_codes(synth.random_seq(self.rng, span + self.k - 1),
                                                 self.k)]
            else:
                seg = r.codes[r.pos:r.pos + span]
                if seg.size < span:  # the walk reached the strand's end: wrap
                    r.pos = 0
                    seg = r.codes[:span]
                r.pos += span
                lv = self.means[seg]
            r.pending = np.concatenate([r.pending, synth.simulate_signal(self.rng, lv, block)])
        out, r.pending = r.pending[:n], r.pending[n:]
        return out

    def _advance(self, chans, now):
        """Channel state at sample `now`: ends reads, starts new ones."""
        for ch, st in chans.items():
            r = st["read"]
            if r is not None and now >= r.end:
                self.finished.append((ch, r))
                st["read"] = None
                st["next_start"] = r.end + int(self.rng.exponential(self.gap_s) * self.fs)
            if st["read"] is None and now >= st["next_start"]:
                st["number"] += 1
                st["read"] = self._new_read(st["number"], st["next_start"])
                self.reads[(ch, st["number"])] = st["read"]

    # ---- service methods ---------------------------------------------------
    def get_data_types(self, request, context):
        P = mk.protocol()
        DT = P.GetDataTypesResponse.DataType
        return P.GetDataTypesResponse(
            uncalibrated_signal=DT(type=P.SIGNED_INTEGER, big_endian=False, size=2),
            calibrated_signal=DT(type=P.FLOATING_POINT, big_endian=False, size=4),
            bias_voltages=DT(type=P.SIGNED_INTEGER, big_endian=False, size=2))

    def get_live_reads(self, request_iterator, context):
        import grpc
        P = mk.protocol()
        first = next(request_iterator, None)
        if first is None or first.WhichOneof("request") != "setup":
            context.abort(grpc.StatusCode.INVALID_ARGUMENT, "the first request must be a setup")
        setup = first.setup
        lo, hi = setup.first_channel, setup.last_channel
        if not (1 <= lo <= hi <= self.n_channels):
            context.abort(grpc.StatusCode.INVALID_ARGUMENT, "bad channel range")
        if setup.raw_data_type != P.RAW_DATA_TYPES["calibrated"]:
            context.abort(grpc.StatusCode.UNIMPLEMENTED, "only calibrated signal is simulated")
        chans = {ch: {"read": None, "number": 0,
                      "next_start": int(self.rng.exponential(self.gap_s) * self.fs)}
                 for ch in range(lo, hi + 1)}
        replies = []
        state = {"now": 0, "open": True}
        self.n_streams += 1

        def actions():
            try:
                for req in request_iterator:
                    if req.WhichOneof("request") != "actions":
                        continue
                    with self.lock:
                        for a in req.actions.actions:
                            replies.append(self._enact(chans, a, state["now"]))
            except Exception:
                pass
            finally:
                state["open"] = False

        threading.Thread(target=actions, daemon=True).start()
        t0 = time.perf_counter()
        while context.is_active():
            time.sleep(self.interval_s)
            now = int((time.perf_counter() - t0) * self.speed * self.fs)
            if self.duration_s is not None and now > self.duration_s * self.fs:
                break
            msg = P.GetLiveReadsResponse(samples_since_start=now, seconds_since_start=now / self.fs)
            with self.lock:
                state["now"] = now
                self._advance(chans, now)
                for ch, st in chans.items():
                    r = st["read"]
                    if r is None or r.stopped:
                        continue
                    upto = min(now, r.end)
                    if upto <= r.sent:
                        continue
                    sig = self._signal(r, upto - r.sent)
                    rd = msg.channels[ch]
                    rd.id, rd.number = r.id, r.number
                    rd.start_sample, rd.chunk_start_sample = r.start, r.sent
                    rd.chunk_length = upto - r.sent
                    rd.raw_data = sig.astype("<f4").tobytes()
                    rd.chunk_classifications.append(mk.CLASS_STRAND)
                    r.sent = upto
                msg.action_responses.extend(replies)
                replies.clear()
            yield msg

    def _enact(self, chans, a, now):
        P = mk.protocol()
        R = P.GetLiveReadsResponse.ActionResponse
        st = chans.get(a.channel)
        r = None if st is None else st["read"]
        if r is None or a.WhichOneof("read") != "number" or a.number != r.number or now >= r.end:
            return R(action_id=a.action_id, response=P.FAILED_READ_FINISHED)
        kind = a.WhichOneof("action")
        if kind == "unblock":
            r.unblocked = True
            r.end = max(now, r.start + 1)
        elif kind == "stop_further_data":
            r.stopped = True
        return R(action_id=a.action_id, response=P.SUCCESS)

    # ---- results -----------------------------------------------------------
    def summary(self):
        """Per class: reads ended, reads unblocked, samples sequenced."""
        with self.lock:
            out = {}
            for label, want in (("target", True), ("nontarget", False)):
                rs = [r for _, r in self.finished if r.target == want]
                out[label] = {"reads": len(rs), "unblocked": sum(r.unblocked for r in rs),
                              "samples": int(sum(r.end - r.start for r in rs))}
            return out

    def truth(self):
        """{(channel, number): is_target} of every read started."""
        with self.lock:
            return {key: r.target for key, r in self.reads.items()}


def main(argv=None):
    """dtwax-sim-sequencer: serve a simulated flow cell (or a POD5 playback)
    over MinKNOW's data service for dry runs of `python -m dtwax_amd.minknow`."""
    import argparse
    from concurrent import futures

    import grpc

    import dtwax_amd as D

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("model", help="k-mer model (KMER mean sd)")
    ap.add_argument("fasta", help="reference FASTA: targets are drawn from it")
    ap.add_argument("--port", type=int, default=8000)
    ap.add_argument("--channels", type=int, default=512)
    ap.add_argument("--speed", type=float, default=1.0, help="simulated seconds per wall second")
    ap.add_argument("--target-frac", type=float, default=0.1)
    ap.add_argument("--read-mean", type=float, default=20000.0, help="mean read length (samples)")
    ap.add_argument("--playback", help="POD5 file whose reads are served instead of simulated ones")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    means = D.read_kmer_model(a.model)
    k = int(round(np.log(means.size) / np.log(4)))
    sim = SimulatedSequencer(D.read_fasta(a.fasta), means, n_channels=a.channels, read_mean=a.read_mean,
                             target_frac=a.target_frac, speed=a.speed, seed=a.seed, k=k,
                             playback=a.playback)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=8))
    mk.add_data_service(server, sim)
    server.add_insecure_port(f"127.0.0.1:{a.port}")
    server.start()
    print(f"serving {a.channels} channels on 127.0.0.1:{a.port}", flush=True)
    try:
        server.wait_for_termination()
    except KeyboardInterrupt:
        server.stop(0)
    print(sim.summary())


if __name__ == "__main__":
    main()
