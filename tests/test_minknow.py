"""MinKNOW Read Until wire client (dtwax_amd.minknow), on the CPU.

* The protocol declaration equals minknow_api 5.4's data.proto for every
  message and method the client uses (tests/golden/minknow_data_fields.json,
  extracted from the reference's vendored generated module), and
  serialises to the protobuf wire bytes written out by hand from those
  field numbers.
* The read cache, the per-read sample assembly (overlaps, gaps, chunking)
  and the action mapping behave as documented.
* Against the simulated sequencer over a loopback gRPC channel, with a
  stand-in engine that knows each read's class, every decided non-target
  read is unblocked and every target read only has its data stopped.
"""
import json
import os
import struct
import time
from concurrent import futures
from types import SimpleNamespace

import numpy as np
import pytest

import dtwax_amd as D
from dtwax_amd import minknow as mk

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _table(m):
    """The fixture's shape for one of our DescriptorProtos."""
    from google.protobuf import descriptor_pb2
    T = descriptor_pb2.FieldDescriptorProto
    return {
        "fields": [{"number": f.number, "name": f.name, "type": T.Type.Name(f.type),
                    "label": T.Label.Name(f.label), "type_name": f.type_name,
                    "oneof": m.oneof_decl[f.oneof_index].name if f.HasField("oneof_index") else None}
                   for f in m.field],
        "nested": {n.name: _table(n) for n in m.nested_type},
        "enums": {e.name: {v.name: v.number for v in e.value} for e in m.enum_type},
        "map_entry": bool(m.options.map_entry),
    }


def test_protocol_matches_minknow_data_proto():
    ref = json.load(open(os.path.join(GOLDEN, "minknow_data_fields.json")))
    stats, data = mk._file_descriptors()
    assert data.package == ref["package"] and data.syntax == ref["syntax"]
    ours = {m.name: _table(m) for m in data.message_type}
    assert set(ours) == set(ref["messages"])
    for name, want in ref["messages"].items():
        assert ours[name] == want, name
    assert {v.name: v.number for v in stats.enum_type[0].value} == ref["read_end_reason"]
    for meth in data.service[0].method:
        want = ref["methods"][meth.name]
        assert f"{data.package}.{data.service[0].name}" == want["service"]
        assert (meth.input_type, meth.output_type, meth.client_streaming,
                meth.server_streaming) == (want["input"], want["output"],
                                           want["client_streaming"], want["server_streaming"])
    assert mk.METHOD_LIVE_READS == "/minknow_api.data.DataService/get_live_reads"


def _varint(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


def _key(field, wire):
    return _varint(field << 3 | wire)


def _ld(field, payload):
    return _key(field, 2) + _varint(len(payload)) + payload


def test_wire_bytes_by_hand():
    P = mk.protocol()
    R = P.GetLiveReadsRequest
    # an unblock action: action_id(1) "a1", channel(2) 5, number(4) 300, unblock(5){duration(1) 0.1}
    act = R.Action(action_id="a1", channel=5, number=300, unblock=R.UnblockAction(duration=0.1))
    body = (_ld(1, b"a1") + _key(2, 0) + _varint(5) + _key(4, 0) + _varint(300)
            + _ld(5, _key(1, 1) + struct.pack("<d", 0.1)))
    req = R(actions=R.Actions(actions=[act]))
    assert req.SerializeToString() == _ld(2, _ld(2, body))
    stop = R.Action(action_id="b", channel=7, number=2, stop_further_data=R.StopFurtherData())
    assert stop.SerializeToString() == _ld(1, b"b") + _key(2, 0) + _varint(7) + _key(4, 0) + \
        _varint(2) + _ld(6, b"")
    setup = R(setup=R.StreamSetup(first_channel=1, last_channel=512, raw_data_type=2,
                                  sample_minimum_chunk_size=0))
    assert setup.SerializeToString() == _ld(1, _key(1, 0) + _varint(1) + _key(2, 0) +
                                            _varint(512) + _key(3, 0) + _varint(2))
    # a response written by hand: channels(4) map entry {key(1) 12, value(2) ReadData}
    sig = np.array([1.5, -2.0, 3.25], "<f4").tobytes()
    rd = (_ld(1, b"r-9") + _key(2, 0) + _varint(9) + _key(3, 0) + _varint(1000)
          + _key(4, 0) + _varint(1600) + _key(5, 0) + _varint(3)
          + _ld(6, _varint(83)) + _ld(7, sig) + _key(8, 5) + struct.pack("<f", 80.5))
    wire = (_key(1, 0) + _varint(4000) + _key(2, 1) + struct.pack("<d", 1.0)
            + _ld(4, _key(1, 0) + _varint(12) + _ld(2, rd))
            + _ld(5, _ld(1, b"a1") + _key(2, 0) + _varint(1)))
    msg = P.GetLiveReadsResponse.FromString(wire)
    assert msg.samples_since_start == 4000 and msg.seconds_since_start == 1.0
    r = msg.channels[12]
    assert (r.id, r.number, r.start_sample, r.chunk_start_sample, r.chunk_length) == \
        ("r-9", 9, 1000, 1600, 3)
    assert list(r.chunk_classifications) == [83] and r.median_before == 80.5
    assert np.array_equal(np.frombuffer(r.raw_data, "<f4"), [1.5, -2.0, 3.25])
    ar = msg.action_responses[0]
    assert ar.action_id == "a1" and ar.response == P.FAILED_READ_FINISHED
    DT = P.GetDataTypesResponse.DataType
    assert mk.numpy_dtype(DT(type=P.FLOATING_POINT, size=4)) == np.dtype("<f4")
    assert mk.numpy_dtype(DT(type=P.SIGNED_INTEGER, size=2, big_endian=True)) == np.dtype(">i2")


def _chunk(number, start, n, fill=1.0):
    return SimpleNamespace(id=f"r{number}", number=number, start_sample=0,
                           chunk_start_sample=start, chunk_length=n, chunk_classifications=[83],
                           raw_data=np.full(n, fill, "<f4").tobytes())


def test_read_cache():
    c = mk.ReadCache(size=2, merge=False)
    c[1] = _chunk(1, 0, 4)
    c[1] = _chunk(1, 4, 4)
    c[2] = _chunk(4, 0, 4)
    c[1] = _chunk(2, 0, 4)
    assert c.replaced == 1 and c.missed == 1 and len(c) == 2
    c[3] = _chunk(1, 0, 4)  # over size: the oldest (channel 2) is dropped
    assert c.missed == 2
    assert [ch for ch, _ in c.popitems(5, last=True)] == [3, 1]
    assert len(c) == 0
    # merging: a contiguous chunk of the queued read is appended
    m = mk.ReadCache(size=4)
    m[1] = _chunk(1, 100, 3, 1.0)
    m[1] = _chunk(1, 103, 2, 2.0)
    m[1] = _chunk(1, 200, 2, 3.0)  # not contiguous: replaces
    m[2] = _chunk(5, 0, 2, 4.0)
    m[2] = _chunk(5, 2, 1, 5.0)
    assert m.merged == 2 and m.replaced == 1
    got = dict(m.popitems(4, last=False))
    assert got[1].chunk_start_sample == 200 and got[1].chunk_length == 2
    assert np.array_equal(np.frombuffer(got[2].raw_data, "<f4"), [4.0, 4.0, 5.0])
    assert got[2].chunk_length == 3 and got[2].chunk_classifications == [83, 83]


class FakeEngine:
    """submit/poll/release like dtwax_amd.ReadUntil; decides after `need`
    chunks from a class lookup (a stand-in for the GPU classifier)."""

    def __init__(self, chunk_len, need, is_target):
        self.chunk_len, self.need, self.is_target = chunk_len, need, is_target
        self.count, self.out, self.released, self.chunks = {}, [], [], {}

    def submit(self, channel, number, samples):
        assert samples.dtype == np.float32 and samples.size == self.chunk_len
        key = (channel, number)
        self.chunks.setdefault(key, []).append(samples.copy())
        self.count[key] = self.count.get(key, 0) + 1
        if self.count[key] == self.need:
            d = np.zeros(1, D.RU_DECISION_DTYPE)
            d["channel"], d["number"], d["chunks"] = channel, number, self.need
            d["action"] = D.RU_TARGET if self.is_target(key) else D.RU_NONTARGET
            self.out.append(d[0])
        elif self.count[key] < self.need:
            d = np.zeros(1, D.RU_DECISION_DTYPE)
            d["channel"], d["number"], d["chunks"] = channel, number, self.count[key]
            self.out.append(d[0])

    def poll(self, cap=1024, timeout_ms=0):
        if not self.out and timeout_ms > 0:  # the engine's poll waits for a decision
            time.sleep(timeout_ms * 1e-3)
        out, self.out = self.out[:cap], self.out[cap:]
        return np.array(out, D.RU_DECISION_DTYPE)

    def release(self, channel):
        self.released.append(channel)


class FakeClient:
    def __init__(self):
        self.signal_dtype = np.dtype("<f4")
        self.actions = []

    def unblock_read(self, ch, num, duration=0.1):
        self.actions.append(("unblock", ch, num))

    def stop_receiving_read(self, ch, num):
        self.actions.append(("stop", ch, num))


def _rd(number, start, sig):
    return SimpleNamespace(number=number, chunk_start_sample=start,
                           raw_data=np.asarray(sig, "<f4").tobytes(), id=f"r{number}")


def test_sample_assembly_and_actions():
    eng = FakeEngine(chunk_len=10, need=3, is_target=lambda k: k[1] == 1)
    cl = FakeClient()
    loop = mk.AdaptiveSampling(cl, eng, eject="nontarget")
    x = np.arange(100, dtype=np.float32)
    loop.ingest(4, _rd(1, 50, x[0:7]))
    loop.ingest(4, _rd(1, 57, x[7:25]))     # contiguous
    loop.ingest(4, _rd(1, 70, x[20:31]))    # overlaps 5 samples: trimmed
    loop.ingest(4, _rd(1, 60, x[10:12]))    # already seen: dropped
    got = np.concatenate(eng.chunks[(4, 1)])
    assert np.array_equal(got, x[:30])
    assert loop.counts["duplicate_chunks"] == 1
    for d in eng.poll():
        loop.act(d)
    # read 1 reached 3 chunks -> TARGET -> its data is stopped
    assert cl.actions == [("stop", 4, 1)]
    # read 2 on the same channel, with a gap: restarted after the gap
    loop.ingest(4, _rd(2, 500, x[:12]))
    loop.ingest(4, _rd(2, 520, x[40:58]))   # gap of 8 samples
    assert loop.counts["gaps"] == 1 and eng.released == [4]
    assert np.array_equal(np.concatenate(eng.chunks[(4, 2)]), np.concatenate([x[:10], x[40:50]]))
    # a late decision for read 1 (the channel has moved on) is not acted on
    d = np.zeros(1, D.RU_DECISION_DTYPE)[0]
    d["channel"], d["number"], d["action"] = 4, 1, D.RU_NONTARGET
    loop.act(d)
    assert cl.actions == [("stop", 4, 1)]
    loop.ingest(4, _rd(2, 538, x[58:70]))
    for d in eng.poll():
        loop.act(d)
    assert cl.actions[-2:] == [("stop", 4, 2), ("unblock", 4, 2)]
    loop.ingest(4, _rd(2, 550, x[70:90]))   # after the decision: ignored
    assert loop.counts["chunks_after_decision"] == 1
    with pytest.raises(ValueError):
        mk.AdaptiveSampling(cl, eng, eject="both")


@pytest.fixture
def sequencer():
    import grpc
    from dtwax_amd import minknow_sim, synth
    rng = np.random.default_rng(3)
    means, _ = synth.make_model(rng)
    seq = synth.random_seq(rng, 20000)
    sim = minknow_sim.SimulatedSequencer(seq, means, n_channels=48, read_mean=6000, gap_s=0.2,
                                         target_frac=0.5, speed=10.0, interval_s=0.02, seed=4)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=4))
    mk.add_data_service(server, sim)
    port = server.add_insecure_port("127.0.0.1:0")
    server.start()
    channel = grpc.insecure_channel(f"127.0.0.1:{port}")
    yield sim, channel
    channel.close()
    server.stop(0)


def test_loop_against_simulated_sequencer(sequencer):
    sim, channel = sequencer
    client = mk.LiveReadsClient(channel, cache_size=48)
    assert client.signal_dtype == np.dtype("<f4")
    truth = {}
    eng = FakeEngine(chunk_len=1600, need=2, is_target=lambda k: truth[k])

    def lookup(k):
        truth.update(sim.truth())
        return truth[k]

    eng.is_target = lookup
    client.run(first_channel=1, last_channel=48)
    loop = mk.AdaptiveSampling(client, eng, eject="nontarget")
    try:
        loop.run(seconds=4.0)
        # the actions already sent get their responses before the stream ends
        # (on a loaded host the last batches can still be in flight)
        deadline = time.time() + 5.0
        while time.time() < deadline:
            got = sum(sum(v.values()) for v in client.responses.values())
            if client.action_queue.empty() and got >= len(client.sent_actions):
                break
            time.sleep(0.02)
    finally:
        client.stop()
    assert client.error is None
    assert loop.counts["reads"] >= 100 and len(loop.decisions) >= 50
    t = sim.truth()
    for ch, num, a, chunks, _ in loop.decisions:
        assert a == (D.RU_TARGET if t[(ch, num)] else D.RU_NONTARGET)
    resp = client.responses
    n_non = sum(a == D.RU_NONTARGET for _, _, a, _, _ in loop.decisions)
    assert n_non >= 10 and resp["unblock"]["success"] >= n_non - 5
    assert resp["stop_further_data"]["success"] >= len(loop.decisions) - 10
    # the sequencer saw exactly the unblocks the loop sent for non-targets
    unblocked = {(ch, r.number) for ch, r in sim.finished if r.unblocked}
    nontarget_decided = {(ch, num) for ch, num, a, _, _ in loop.decisions if a == D.RU_NONTARGET}
    assert unblocked <= nontarget_decided
    assert all(not t[k] for k in unblocked)
    s = sim.summary()
    assert s["target"]["unblocked"] == 0
    # every decided read reached the engine with the two chunks its decision
    # needs; a chunk that arrived while that decision was on its way may have
    # been submitted too (the real engine answers it STALE, unscored).  Chunk
    # assembly itself is pinned by test_sample_assembly_and_actions.
    for ch, num, a, chunks, _ in loop.decisions:
        assert len(eng.chunks[(ch, num)]) >= 2


def test_setup_errors(sequencer):
    import grpc
    sim, channel = sequencer
    stub = mk.DataServiceStub(channel)
    P = mk.protocol()
    R = P.GetLiveReadsRequest
    call = stub.get_live_reads(iter([R(actions=R.Actions())]))
    with pytest.raises(grpc.RpcError) as e:
        next(call)
    assert e.value.code() == grpc.StatusCode.INVALID_ARGUMENT
    call = stub.get_live_reads(iter([R(setup=R.StreamSetup(first_channel=1, last_channel=999,
                                                           raw_data_type=2))]))
    with pytest.raises(grpc.RpcError) as e:
        next(call)
    assert e.value.code() == grpc.StatusCode.INVALID_ARGUMENT


def test_playback_of_a_pod5_file(tmp_path):
    """The simulated sequencer replays a POD5 file: each read arrives under
    its POD5 read id, as contiguous chunks of its calibrated signal, and the
    truth labels come from playback_truth."""
    import threading
    from dtwax_amd import minknow_sim, synth
    from pod5_helpers import make_read, write_pod5
    rng = np.random.default_rng(8)
    reads = [make_read(rng.integers(-500, 1500, n).astype(np.int16), i) for i, n in
             enumerate([9000, 0, 4000, 12000])]
    write_pod5(tmp_path / "run.pod5", reads, row_samples=2500)
    means, _ = synth.make_model(rng)
    truth = {reads[0]["read_id"]: True}
    sim = minknow_sim.SimulatedSequencer(synth.random_seq(rng, 2000), means, n_channels=3, gap_s=0.05,
                                         speed=40.0, interval_s=0.01, duration_s=6.0, seed=1,
                                         playback=str(tmp_path / "run.pod5"), playback_truth=truth)
    P = mk.protocol()
    R = P.GetLiveReadsRequest
    done = threading.Event()

    def requests():
        yield R(setup=R.StreamSetup(first_channel=1, last_channel=3,
                                    raw_data_type=P.RAW_DATA_TYPES["calibrated"]))
        done.wait(30)

    ctx = SimpleNamespace(is_active=lambda: True, abort=None)
    got = {}  # (channel, read number) -> (read id, chunks); the file's reads cycle
    for msg in sim.get_live_reads(requests(), ctx):
        for ch, rd in msg.channels.items():
            rid, parts = got.setdefault((ch, rd.number), (rd.id, []))
            assert rd.id == rid
            assert rd.chunk_start_sample == rd.start_sample + sum(p.size for p in parts)
            parts.append(np.frombuffer(rd.raw_data, "<f4"))
    done.set()
    want = {r["read_id"]: (r["signal"].astype(np.float32) + np.float32(r["calibration_offset"]))
            * np.float32(r["calibration_scale"]) for r in reads if r["signal"].size}
    assert len(got) >= 3 and {rid for rid, _ in got.values()} <= set(want)
    for rid, parts in got.values():
        x = np.concatenate(parts)
        assert x.size <= want[rid].size
        assert np.array_equal(x, want[rid][:x.size])
    s = sim.summary()
    assert s["target"]["reads"] + s["nontarget"]["reads"] >= 2
    assert all(r.target == (r.id == reads[0]["read_id"]) for r in sim.reads.values())
