"""Read Until over the MinKNOW wire protocol with the GPU engine: the
simulated sequencer streams live reads over loopback gRPC, the client feeds
the dtwax_ru engine, and its decisions go back as actions.  The engine's
decisions must match the reads' classes (target = drawn from the reference)
as well as its calibration allows, and the sequencer must enact them:
non-targets unblocked, targets sequenced on."""
from concurrent import futures

import numpy as np
import pytest

import dtwax_amd as D
from dtwax_amd import minknow as mk
from dtwax_amd import minknow_sim, ru_sim, synth

pytestmark = pytest.mark.gpu


def test_adaptive_sampling_over_grpc():
    import grpc
    rng = np.random.default_rng(11)
    means, _ = synth.make_model(rng)
    seq, levels = synth.build_reference(rng, 200_000, means, ref_norm="meanstd")
    ctx = D.Context(0)
    ctx.upload_reference(levels)
    src = ru_sim.ReadSource(seq, means, np.random.default_rng(12))
    thr, cal = ru_sim.calibrate_threshold_per_row(ctx, src, 1600, 4, n_each=48,
                                                  norm=D.NORM_MEANSTD)
    print(f"threshold_per_row {thr:.4f}, calibration accuracy {cal:.3f}", flush=True)
    assert cal >= 0.95
    sim = minknow_sim.SimulatedSequencer(seq, means, n_channels=64, read_mean=12000, gap_s=0.3,
                                         target_frac=0.3, speed=4.0, interval_s=0.05, seed=13)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=4))
    mk.add_data_service(server, sim)
    port = server.add_insecure_port("127.0.0.1:0")
    server.start()
    channel = grpc.insecure_channel(f"127.0.0.1:{port}")
    ru = D.ReadUntil(ctx, max_reads=64, chunk_len=1600, max_chunks=4, norm=D.NORM_MEANSTD,
                     threshold_per_row=thr)
    client = mk.LiveReadsClient(channel, cache_size=64)
    try:
        client.run(first_channel=1, last_channel=64)
        ru.start()
        loop = mk.AdaptiveSampling(client, ru, eject="nontarget")
        loop.run(seconds=8.0)
    finally:
        client.stop()
        ru.stop()
        ru.close()
        channel.close()
        server.stop(0)
        ctx.close()
    s = loop.summary()
    print(s, sim.summary(), flush=True)
    assert client.error is None
    truth = sim.truth()
    dec = [(truth[(ch, num)], a) for ch, num, a, _, _ in loop.decisions
           if a in (D.RU_TARGET, D.RU_NONTARGET)]
    assert len(dec) >= 100
    tgt = np.array([t for t, _ in dec])
    said = np.array([a == D.RU_TARGET for _, a in dec])
    acc = ((said & tgt).sum() / max(tgt.sum(), 1) + (~said & ~tgt).sum() / max((~tgt).sum(), 1)) / 2
    print(f"decisions {len(dec)}, balanced accuracy {acc:.3f}", flush=True)
    assert acc >= 0.9
    # the sequencer enacted them: unblocked reads are exactly the ones the
    # engine called non-target (a few may still be in flight at the end)
    unblocked = {(ch, r.number) for ch, r in sim.finished if r.unblocked}
    called_non = {(ch, num) for ch, num, a, _, _ in loop.decisions if a == D.RU_NONTARGET}
    assert unblocked <= called_non and len(unblocked) >= 0.8 * len(called_non)
    assert s["responses"]["unblock"].get("success", 0) >= 0.8 * len(called_non)
    # enrichment: targets hold a larger share of the sequenced samples than of the reads
    sm = sim.summary()
    frac_reads = sm["target"]["reads"] / max(sm["target"]["reads"] + sm["nontarget"]["reads"], 1)
    frac_samples = sm["target"]["samples"] / max(sm["target"]["samples"] + sm["nontarget"]["samples"], 1)
    assert frac_samples > frac_reads
