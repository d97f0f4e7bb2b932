"""MinKNOW Read Until wire client: live raw-signal chunks from a sequencer's
data service in, the Read Until engine's decisions out as unblock /
stop-further-data actions.

The reference joins its classifier to MinKNOW in one direction only:
src/ReadUntil/ont_readuntil.py:29-59 pulls chunks through read_until's
ReadUntilClient, unblocks every read unconditionally and writes raw bytes
into a FIFO for the CUDA process, whose scores are printed and never acted
on.  Here the decisions close the loop.

Wire protocol
    minknow_api.data.DataService as minknow_api 5.4's data.proto defines it
    (the reference vendors the generated module,
    src/ReadUntil/read_until_env/lib/python3.6/site-packages/minknow_api/data_pb2.py):
        get_live_reads(stream GetLiveReadsRequest) -> stream GetLiveReadsResponse
        get_data_types(GetDataTypesRequest) -> GetDataTypesResponse
    `protocol()` declares the messages these two methods carry with the same
    package, names, field numbers, types, labels and oneofs (proto3), so the
    client talks to MinKNOW and to anything built from minknow_api without
    that package.  Connection set-up (TLS roots, tokens, the manager's port
    lookup) is the caller's: pass any grpc.Channel.

Client (`LiveReadsClient`)
    Follows read_until 2.1's ReadUntilClient (read_until/base.py:166-567):
    run() opens the stream with a StreamSetup (channel range, calibrated
    signal, minimum chunk size), reset() ends it, get_read_chunks() pops the
    most recent chunk per channel, unblock_read() / stop_receiving_read()
    queue actions (uuid action ids, sent in batches), signal_dtype comes from
    get_data_types, and action responses are counted per action type.  The
    one-chunk mode (stop data after the first chunk) is optional and off by
    default: the engine classifies from several chunks.

Loop (`AdaptiveSampling`)
    Chunks -> per-read sample streams -> engine chunks of chunk_len samples
    (dtwax_ru_submit) -> decisions (dtwax_ru_poll) -> actions: the ejected
    class is unblocked (and its data stopped), the kept class has its data
    stopped and sequences on.
"""
from __future__ import annotations

import argparse
import collections
import queue
import threading
import time
import uuid
from types import SimpleNamespace

import numpy as np

PACKAGE = "minknow_api.data"
SERVICE = PACKAGE + ".DataService"
METHOD_LIVE_READS = f"/{SERVICE}/get_live_reads"
METHOD_DATA_TYPES = f"/{SERVICE}/get_data_types"

# MinKNOW read classifications (read_until/base.py:34-50): strand-like chunks
CLASS_STRAND, CLASS_ADAPTER = 83, 65
STRAND_CLASSES = (CLASS_STRAND, CLASS_ADAPTER)

_proto = None
_proto_lock = threading.Lock()


def _file_descriptors():
    """(statistics.proto subset, data.proto subset) as FileDescriptorProtos.

    Field tables (number name type [oneof]) of minknow_api 5.4's data.proto:
      GetLiveReadsRequest: 1 setup StreamSetup [request], 2 actions Actions [request]
        StreamSetup: 1 first_channel u32, 2 last_channel u32,
          3 raw_data_type RawDataType, 4 sample_minimum_chunk_size u64,
          5 max_unblock_read_length_samples u64 [max_unblock_read_length],
          6 max_unblock_read_length_seconds double [max_unblock_read_length],
          7 accepted_first_chunk_classifications repeated i32
        Actions: 2 actions repeated Action
        Action: 1 action_id string, 2 channel u32, 3 id string [read],
          4 number u32 [read], 5 unblock UnblockAction [action],
          6 stop_further_data StopFurtherData [action]
        UnblockAction: 1 duration double;  StopFurtherData: (empty)
        RawDataType: KEEP_LAST 0, NONE 1, CALIBRATED 2, UNCALIBRATED 3
      GetLiveReadsResponse: 1 samples_since_start u64, 2 seconds_since_start
        double, 4 channels map<u32, ReadData>, 5 action_responses repeated
        ActionResponse
        ReadData: 1 id string, 2 number u32, 3 start_sample u64,
          4 chunk_start_sample u64, 5 chunk_length u64,
          6 chunk_classifications repeated i32, 7 raw_data bytes,
          8 median_before float, 9 median float,
          10 previous_read_classification i32,
          11 previous_read_end_reason minknow_api.statistics.ReadEndReason
        ActionResponse: 1 action_id string, 2 response Response
          (SUCCESS 0, FAILED_READ_FINISHED 1, FAILED_READ_TOO_LONG 2)
      GetDataTypesRequest: (empty)
      GetDataTypesResponse: 1 uncalibrated_signal, 2 calibrated_signal,
        3 bias_voltages: DataType {1 type Type, 2 big_endian bool, 3 size u32;
        Type: SIGNED_INTEGER 0, UNSIGNED_INTEGER 1, FLOATING_POINT 2}
    """
    from google.protobuf import descriptor_pb2 as dp

    F = dp.FieldDescriptorProto

    def msg(name, fields=(), nested=(), enums=(), oneofs=(), map_entry=False):
        m = dp.DescriptorProto(name=name)
        for o in oneofs:
            m.oneof_decl.add(name=o)
        for number, fname, ftype, extra in fields:
            f = m.field.add(name=fname, number=number, type=ftype,
                            label=extra.get("label", F.LABEL_OPTIONAL), json_name=_camel(fname))
            if "type_name" in extra:
                f.type_name = extra["type_name"]
            if "oneof" in extra:
                f.oneof_index = list(oneofs).index(extra["oneof"])
        m.nested_type.extend(nested)
        m.enum_type.extend(enums)
        if map_entry:
            m.options.map_entry = True
        return m

    def enum(name, values):
        e = dp.EnumDescriptorProto(name=name)
        for n, v in values:
            e.value.add(name=n, number=v)
        return e

    REP = {"label": F.LABEL_REPEATED}
    req = ".minknow_api.data.GetLiveReadsRequest."
    resp = ".minknow_api.data.GetLiveReadsResponse."
    dtr = ".minknow_api.data.GetDataTypesResponse."

    stats = dp.FileDescriptorProto(name="minknow_api/statistics.proto",
                                   package="minknow_api.statistics", syntax="proto3")
    stats.enum_type.append(enum("ReadEndReason", [
        ("All", 0), ("Unknown", 1), ("Partial", 2), ("MuxChange", 3), ("UnblockMuxChange", 4),
        ("SignalPositive", 5), ("SignalNegative", 6), ("DataServiceUnblockMuxChange", 7)]))

    request = msg(
        "GetLiveReadsRequest",
        fields=[(1, "setup", F.TYPE_MESSAGE, {"type_name": req + "StreamSetup", "oneof": "request"}),
                (2, "actions", F.TYPE_MESSAGE, {"type_name": req + "Actions", "oneof": "request"})],
        oneofs=["request"],
        nested=[
            msg("UnblockAction", [(1, "duration", F.TYPE_DOUBLE, {})]),
            msg("StopFurtherData"),
            msg("Action",
                [(1, "action_id", F.TYPE_STRING, {}), (2, "channel", F.TYPE_UINT32, {}),
                 (3, "id", F.TYPE_STRING, {"oneof": "read"}),
                 (4, "number", F.TYPE_UINT32, {"oneof": "read"}),
                 (5, "unblock", F.TYPE_MESSAGE, {"type_name": req + "UnblockAction",
                                                 "oneof": "action"}),
                 (6, "stop_further_data", F.TYPE_MESSAGE, {"type_name": req + "StopFurtherData",
                                                           "oneof": "action"})],
                oneofs=["read", "action"]),
            msg("StreamSetup",
                [(1, "first_channel", F.TYPE_UINT32, {}), (2, "last_channel", F.TYPE_UINT32, {}),
                 (3, "raw_data_type", F.TYPE_ENUM, {"type_name": req + "RawDataType"}),
                 (4, "sample_minimum_chunk_size", F.TYPE_UINT64, {}),
                 (5, "max_unblock_read_length_samples", F.TYPE_UINT64,
                  {"oneof": "max_unblock_read_length"}),
                 (6, "max_unblock_read_length_seconds", F.TYPE_DOUBLE,
                  {"oneof": "max_unblock_read_length"}),
                 (7, "accepted_first_chunk_classifications", F.TYPE_INT32, REP)],
                oneofs=["max_unblock_read_length"]),
            msg("Actions", [(2, "actions", F.TYPE_MESSAGE, {"type_name": req + "Action", **REP})]),
        ],
        enums=[enum("RawDataType", [("KEEP_LAST", 0), ("NONE", 1), ("CALIBRATED", 2),
                                    ("UNCALIBRATED", 3)])])

    response = msg(
        "GetLiveReadsResponse",
        fields=[(1, "samples_since_start", F.TYPE_UINT64, {}),
                (2, "seconds_since_start", F.TYPE_DOUBLE, {}),
                (4, "channels", F.TYPE_MESSAGE, {"type_name": resp + "ChannelsEntry", **REP}),
                (5, "action_responses", F.TYPE_MESSAGE, {"type_name": resp + "ActionResponse",
                                                         **REP})],
        nested=[
            msg("ReadData",
                [(1, "id", F.TYPE_STRING, {}), (2, "number", F.TYPE_UINT32, {}),
                 (3, "start_sample", F.TYPE_UINT64, {}),
                 (4, "chunk_start_sample", F.TYPE_UINT64, {}),
                 (5, "chunk_length", F.TYPE_UINT64, {}),
                 (6, "chunk_classifications", F.TYPE_INT32, REP),
                 (7, "raw_data", F.TYPE_BYTES, {}), (8, "median_before", F.TYPE_FLOAT, {}),
                 (9, "median", F.TYPE_FLOAT, {}),
                 (10, "previous_read_classification", F.TYPE_INT32, {}),
                 (11, "previous_read_end_reason", F.TYPE_ENUM,
                  {"type_name": ".minknow_api.statistics.ReadEndReason"})]),
            msg("ActionResponse",
                [(1, "action_id", F.TYPE_STRING, {}),
                 (2, "response", F.TYPE_ENUM, {"type_name": resp + "ActionResponse.Response"})],
                enums=[enum("Response", [("SUCCESS", 0), ("FAILED_READ_FINISHED", 1),
                                         ("FAILED_READ_TOO_LONG", 2)])]),
            msg("ChannelsEntry",
                [(1, "key", F.TYPE_UINT32, {}),
                 (2, "value", F.TYPE_MESSAGE, {"type_name": resp + "ReadData"})],
                map_entry=True),
        ])

    data_types = msg(
        "GetDataTypesResponse",
        fields=[(1, "uncalibrated_signal", F.TYPE_MESSAGE, {"type_name": dtr + "DataType"}),
                (2, "calibrated_signal", F.TYPE_MESSAGE, {"type_name": dtr + "DataType"}),
                (3, "bias_voltages", F.TYPE_MESSAGE, {"type_name": dtr + "DataType"})],
        nested=[msg("DataType",
                    [(1, "type", F.TYPE_ENUM, {"type_name": dtr + "DataType.Type"}),
                     (2, "big_endian", F.TYPE_BOOL, {}), (3, "size", F.TYPE_UINT32, {})],
                    enums=[enum("Type", [("SIGNED_INTEGER", 0), ("UNSIGNED_INTEGER", 1),
                                         ("FLOATING_POINT", 2)])])])

    data = dp.FileDescriptorProto(name="minknow_api/data.proto", package=PACKAGE,
                                  syntax="proto3", dependency=["minknow_api/statistics.proto"])
    data.message_type.extend([msg("GetDataTypesRequest"), data_types, request, response])
    svc = data.service.add(name="DataService")
    svc.method.add(name="get_live_reads", input_type=".minknow_api.data.GetLiveReadsRequest",
                   output_type=".minknow_api.data.GetLiveReadsResponse",
                   client_streaming=True, server_streaming=True)
    svc.method.add(name="get_data_types", input_type=".minknow_api.data.GetDataTypesRequest",
                   output_type=".minknow_api.data.GetDataTypesResponse")
    return stats, data


def _camel(name):
    head, *rest = name.split("_")
    return head + "".join(w[:1].upper() + w[1:] for w in rest)


def protocol():
    """Message classes of the data service subset (built once, in a private
    descriptor pool so they never clash with an installed minknow_api)."""
    global _proto
    with _proto_lock:
        if _proto is None:
            from google.protobuf import descriptor_pool, message_factory
            stats, data = _file_descriptors()
            pool = descriptor_pool.DescriptorPool()
            pool.Add(stats)
            pool.Add(data)

            def cls(name):
                return message_factory.GetMessageClass(pool.FindMessageTypeByName(f"{PACKAGE}.{name}"))

            req = cls("GetLiveReadsRequest")
            resp = cls("GetLiveReadsResponse")
            _proto = SimpleNamespace(
                pool=pool, files=(stats, data),
                GetLiveReadsRequest=req, GetLiveReadsResponse=resp,
                GetDataTypesRequest=cls("GetDataTypesRequest"),
                GetDataTypesResponse=cls("GetDataTypesResponse"),
                RAW_DATA_TYPES={"keep_last": 0, "none": 1, "calibrated": 2, "uncalibrated": 3},
                SUCCESS=0, FAILED_READ_FINISHED=1, FAILED_READ_TOO_LONG=2,
                RESPONSE_NAMES={0: "success", 1: "failed_read_finished", 2: "failed_read_too_long"},
                SIGNED_INTEGER=0, UNSIGNED_INTEGER=1, FLOATING_POINT=2)
    return _proto


def numpy_dtype(data_type):
    """numpy dtype of a GetDataTypesResponse.DataType (type, size, endianness)."""
    P = protocol()
    kind = {P.SIGNED_INTEGER: "i", P.UNSIGNED_INTEGER: "u", P.FLOATING_POINT: "f"}[data_type.type]
    return np.dtype(f"{'>' if data_type.big_endian else '<'}{kind}{data_type.size}")


class DataServiceStub:
    """The two data-service methods on a grpc.Channel."""

    def __init__(self, channel):
        P = protocol()
        self.get_live_reads = channel.stream_stream(
            METHOD_LIVE_READS, request_serializer=P.GetLiveReadsRequest.SerializeToString,
            response_deserializer=P.GetLiveReadsResponse.FromString)
        self.get_data_types = channel.unary_unary(
            METHOD_DATA_TYPES, request_serializer=P.GetDataTypesRequest.SerializeToString,
            response_deserializer=P.GetDataTypesResponse.FromString)


def add_data_service(server, servicer):
    """Registers servicer.get_live_reads(request_iterator, context) and
    servicer.get_data_types(request, context) on a grpc.Server."""
    import grpc
    P = protocol()
    handlers = {
        "get_live_reads": grpc.stream_stream_rpc_method_handler(
            servicer.get_live_reads, request_deserializer=P.GetLiveReadsRequest.FromString,
            response_serializer=P.GetLiveReadsResponse.SerializeToString),
        "get_data_types": grpc.unary_unary_rpc_method_handler(
            servicer.get_data_types, request_deserializer=P.GetDataTypesRequest.FromString,
            response_serializer=P.GetDataTypesResponse.SerializeToString),
    }
    server.add_generic_rpc_handlers((grpc.method_handlers_generic_handler(SERVICE, handlers),))


class Chunk:
    """A read's queued samples (the ReadData fields the loop uses); several
    contiguous ReadData of one read merge into one Chunk."""
    __slots__ = ("id", "number", "start_sample", "chunk_start_sample", "chunk_length",
                 "chunk_classifications", "raw_data")

    def __init__(self, rd):
        self.id, self.number = rd.id, rd.number
        self.start_sample, self.chunk_start_sample = rd.start_sample, rd.chunk_start_sample
        self.chunk_length = rd.chunk_length
        self.chunk_classifications = list(rd.chunk_classifications)
        self.raw_data = bytes(rd.raw_data)


class ReadCache:
    """Latest data per channel, in arrival order (read_until's ReadCache,
    read_until/base.py:53-160).  A chunk of the read already queued on its
    channel is appended when it continues it (merge=True; the engine scores
    several chunks per read, so none may be lost while the loop is busy) and
    replaces it otherwise; `replaced` counts replacements by the same read,
    `missed` by a new read."""

    def __init__(self, size=512, merge=True):
        self.size = size
        self.merge = merge
        self._d = collections.OrderedDict()
        self._lock = threading.Lock()
        self.replaced = 0
        self.missed = 0
        self.merged = 0

    def __setitem__(self, channel, read):
        with self._lock:
            old = self._d.pop(channel, None)
            if old is not None:
                if old.number == read.number:
                    if self.merge and isinstance(old, Chunk) and \
                            read.chunk_start_sample == old.chunk_start_sample + old.chunk_length:
                        old.raw_data += bytes(read.raw_data)
                        old.chunk_length += read.chunk_length
                        old.chunk_classifications.extend(read.chunk_classifications)
                        self._d[channel] = old
                        self.merged += 1
                        return
                    self.replaced += 1
                else:
                    self.missed += 1
            self._d[channel] = Chunk(read) if self.merge else read
            while len(self._d) > self.size:
                self._d.popitem(last=False)
                self.missed += 1

    def __len__(self):
        return len(self._d)

    def popitems(self, items=1, last=True):
        out = []
        with self._lock:
            while self._d and len(out) < items:
                out.append(self._d.popitem(last=last))
        return out


class LiveReadsClient:
    """Read Until stream client (see the module docstring)."""

    def __init__(self, channel, cache_size=512, one_chunk=False, strand_classes=None,
                 action_batch=1000, merge=True):
        P = protocol()
        self.stub = DataServiceStub(channel)
        self.signal_dtype = numpy_dtype(self.stub.get_data_types(P.GetDataTypesRequest())
                                        .calibrated_signal)
        self.cache_size = cache_size
        self.merge = merge
        self.one_chunk = one_chunk
        self.strand_classes = None if strand_classes is None else set(strand_classes)
        self.action_batch = action_batch
        self._thread = None
        self._call = None
        self.error = None
        self.reset()

    # ---- stream life cycle ------------------------------------------------
    def run(self, first_channel=1, last_channel=512, raw_data_type="calibrated",
            min_chunk_size=0, max_unblock_read_length_samples=0):
        P = protocol()
        setup = P.GetLiveReadsRequest.StreamSetup(
            first_channel=first_channel, last_channel=last_channel,
            raw_data_type=P.RAW_DATA_TYPES[raw_data_type],
            sample_minimum_chunk_size=min_chunk_size)
        if max_unblock_read_length_samples:
            setup.max_unblock_read_length_samples = max_unblock_read_length_samples
        self.running.set()
        self._thread = threading.Thread(target=self._run, args=(P.GetLiveReadsRequest(setup=setup),),
                                        name="dtwax-live-reads", daemon=True)
        self._thread.start()

    def stop(self, timeout=5.0):
        """Ends the stream (if any); queues and counters are kept."""
        if self._thread is not None:
            self.running.clear()
            call = self._call
            if call is not None:
                call.cancel()
            self._thread.join(timeout)
        self._thread = None
        self._call = None

    def reset(self, timeout=5.0):
        """Ends the stream (if any) and clears queues and counters."""
        self.stop(timeout)
        self.running = threading.Event()
        self.action_queue = queue.Queue()
        self.data_queue = ReadCache(self.cache_size, self.merge)
        self.sent_actions = {}
        self.responses = collections.defaultdict(collections.Counter)
        self.seen_reads = set()
        self.n_messages = 0
        self.n_chunks = 0

    @property
    def is_running(self):
        return self.running.is_set()

    def _requests(self, setup):
        yield setup
        P = protocol()
        while self.running.is_set():
            try:
                first = self.action_queue.get(timeout=0.005)
            except queue.Empty:
                continue
            acts = [first]
            while len(acts) < self.action_batch:
                try:
                    acts.append(self.action_queue.get_nowait())
                except queue.Empty:
                    break
            yield P.GetLiveReadsRequest(actions=P.GetLiveReadsRequest.Actions(actions=acts))

    def _run(self, setup):
        import grpc
        try:
            self._call = self.stub.get_live_reads(self._requests(setup))
            for msg in self._call:
                if not self.running.is_set():
                    break
                self._on_message(msg)
        except grpc.RpcError as e:
            if e.code() != grpc.StatusCode.CANCELLED:
                self.error = e
        finally:
            self.running.clear()

    def _on_message(self, msg):
        P = protocol()
        self.n_messages += 1
        for r in msg.action_responses:
            kind = self.sent_actions.get(r.action_id, "unknown")
            self.responses[kind][P.RESPONSE_NAMES.get(r.response, str(r.response))] += 1
        for channel, read in msg.channels.items():
            self.n_chunks += 1
            if self.one_chunk:
                if read.id in self.seen_reads:
                    continue  # a stop request was not enacted in time
                self.stop_receiving_read(channel, read.number)
            self.seen_reads.add(read.id)
            if self.strand_classes is not None and \
                    not any(c in self.strand_classes for c in read.chunk_classifications):
                continue
            self.data_queue[channel] = read

    # ---- data and actions -------------------------------------------------
    def get_read_chunks(self, batch_size=1, last=True):
        """Pops up to batch_size (channel, ReadData) pairs, most recent first."""
        return self.data_queue.popitems(items=batch_size, last=last)

    def unblock_read(self, channel, number, duration=0.1):
        P = protocol()
        self._put(channel, number, "unblock",
                  unblock=P.GetLiveReadsRequest.UnblockAction(duration=duration))

    def stop_receiving_read(self, channel, number):
        P = protocol()
        self._put(channel, number, "stop_further_data",
                  stop_further_data=P.GetLiveReadsRequest.StopFurtherData())

    def _put(self, channel, number, kind, **action):
        P = protocol()
        aid = str(uuid.uuid4())
        self.sent_actions[aid] = kind
        self.action_queue.put(P.GetLiveReadsRequest.Action(action_id=aid, channel=channel,
                                                           number=number, **action))


class _ReadStream:
    __slots__ = ("number", "next_sample", "parts", "n", "decided", "submitted", "t_last")

    def __init__(self, number, start):
        self.number = number
        self.next_sample = start
        self.parts = []
        self.n = 0
        self.decided = False
        self.submitted = 0
        self.t_last = 0.0


class AdaptiveSampling:
    """Client chunks -> engine -> actions (see the module docstring).

    engine: a started dtwax_amd.ReadUntil (submit / poll / release).
    eject: "nontarget" (enrich the reference's reads: unblock the rest) or
        "target" (deplete them).
    Per channel the read's samples are kept in order by chunk_start_sample;
    an overlap is trimmed, a gap restarts the read's classification from the
    samples after it (the DP needs contiguous signal).
    """

    def __init__(self, client, engine, eject="nontarget", unblock_duration=0.1,
                 batch_size=512):
        from . import RU_NONTARGET, RU_TARGET
        if eject not in ("nontarget", "target"):
            raise ValueError("eject must be 'nontarget' or 'target'")
        self.client = client
        self.engine = engine
        self.L = engine.chunk_len
        self.eject_action = RU_NONTARGET if eject == "nontarget" else RU_TARGET
        self.keep_action = RU_TARGET if eject == "nontarget" else RU_NONTARGET
        self.unblock_duration = unblock_duration
        self.batch_size = batch_size
        self.reads = {}
        self.counts = collections.Counter()
        self.latency_s = []  # last engine chunk submitted -> action queued
        self.decisions = []  # (channel, number, action, chunks, score)

    def ingest(self, channel, read, now=None):
        now = time.perf_counter() if now is None else now
        st = self.reads.get(channel)
        if st is None or st.number != read.number:
            st = _ReadStream(read.number, read.chunk_start_sample)
            self.reads[channel] = st
            self.counts["reads"] += 1
        if st.decided:
            self.counts["chunks_after_decision"] += 1
            return
        sig = np.frombuffer(read.raw_data, self.client.signal_dtype).astype(np.float32)
        start = read.chunk_start_sample
        if start > st.next_sample:  # gap: restart from here
            self.counts["gaps"] += 1
            self.engine.release(channel)
            st.parts, st.n, st.submitted = [], 0, 0
            st.next_sample = start
        skip = st.next_sample - start
        if skip >= sig.size:
            self.counts["duplicate_chunks"] += 1
            return
        if skip > 0:
            sig = sig[skip:]
        st.parts.append(sig)
        st.n += sig.size
        st.next_sample = start + skip + sig.size
        if st.n >= self.L:
            buf = np.concatenate(st.parts)
            k = 0
            while buf.size - k >= self.L:
                self.engine.submit(channel, st.number, buf[k:k + self.L])
                st.submitted += 1
                k += self.L
            st.parts = [buf[k:]] if k < buf.size else []
            st.n = buf.size - k
            st.t_last = now

    def act(self, d, now=None):
        from . import RU_ACTIONS
        now = time.perf_counter() if now is None else now
        ch, num, a = int(d["channel"]), int(d["number"]), int(d["action"])
        self.counts[RU_ACTIONS.get(a, str(a))] += 1
        st = self.reads.get(ch)
        if st is None or st.number != num or st.decided:
            return
        if a == self.eject_action:  # as ont_readuntil.py:45-46: stop data, then unblock
            self.client.stop_receiving_read(ch, num)
            self.client.unblock_read(ch, num, self.unblock_duration)
        elif a == self.keep_action:
            self.client.stop_receiving_read(ch, num)
        elif a != 0:  # no slot: the read goes unclassified; stop its data
            self.client.stop_receiving_read(ch, num)
        else:
            return
        st.decided = True
        st.parts = []
        self.latency_s.append(now - st.t_last)
        self.decisions.append((ch, num, a, int(d["chunks"]), float(np.min(d["score"]))))

    def step(self, timeout_ms=2):
        for ch, read in self.client.get_read_chunks(self.batch_size, last=True):
            self.ingest(ch, read)
        for d in self.engine.poll(4096, timeout_ms=timeout_ms):
            self.act(d)

    def run(self, seconds=None, until=None):
        """Steps while the client's stream runs (and until `seconds` pass or
        `until()` is true)."""
        t0 = time.perf_counter()
        while self.client.is_running:
            self.step()
            if seconds is not None and time.perf_counter() - t0 >= seconds:
                break
            if until is not None and until():
                break
        # answer what is still in flight
        for d in self.engine.poll(4096, timeout_ms=50):
            self.act(d)

    def summary(self):
        lat = np.array(self.latency_s) if self.latency_s else np.zeros(1)
        return {"counts": dict(self.counts), "decided": len(self.decisions),
                "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
                "latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
                "responses": {k: dict(v) for k, v in self.client.responses.items()}}


def main(argv=None):
    """dtwax-readuntil: classify live reads from MinKNOW and act on them."""
    import grpc

    import dtwax_amd as D

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("model", help="k-mer model (KMER mean sd)")
    ap.add_argument("fasta", help="reference FASTA")
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=8000)
    ap.add_argument("--tls-roots", help="PEM root certificates (secure channel)")
    ap.add_argument("--first-channel", type=int, default=1)
    ap.add_argument("--last-channel", type=int, default=512)
    ap.add_argument("--chunk-len", type=int, default=1600)
    ap.add_argument("--max-chunks", type=int, default=4)
    ap.add_argument("--threshold-per-row", type=float, required=True)
    ap.add_argument("--eject", choices=("nontarget", "target"), default="nontarget")
    ap.add_argument("--seconds", type=float, default=None)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--precision", choices=("f16", "f16_fused", "f16_renorm", "f32"), default="f16",
                    help="arithmetic form of the serving engine (f16: the reference's half2 ops; "
                         "f16_fused: ~25%% more chunks per second under a full burst)")
    a = ap.parse_args(argv)

    target = f"{a.host}:{a.port}"
    if a.tls_roots:
        creds = grpc.ssl_channel_credentials(open(a.tls_roots, "rb").read())
        channel = grpc.secure_channel(target, creds)
    else:
        channel = grpc.insecure_channel(target)
    ctx = D.Context(a.device)
    means = D.read_kmer_model(a.model)
    k = int(round(np.log(means.size) / np.log(4)))
    levels, _ = D.squiggle_build(D.read_fasta(a.fasta), means, k=k)
    ctx.upload_reference(levels)
    n = a.last_channel - a.first_channel + 1
    ru = D.ReadUntil(ctx, max_reads=n, chunk_len=a.chunk_len, max_chunks=a.max_chunks, precision=a.precision,
                     norm=D.NORM_MEANSTD, threshold_per_row=a.threshold_per_row)
    client = LiveReadsClient(channel, cache_size=n)
    client.run(first_channel=a.first_channel, last_channel=a.last_channel)
    ru.start()
    loop = AdaptiveSampling(client, ru, eject=a.eject)
    try:
        loop.run(seconds=a.seconds)
    except KeyboardInterrupt:
        pass
    finally:
        client.stop()
        ru.stop()
        print(loop.summary())
        ru.close()
        ctx.close()


if __name__ == "__main__":
    main()
