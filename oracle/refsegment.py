"""refsegment — CPU restatement of the Segment container's byte format (test infrastructure).

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the library's
Segment serializer and parser; nothing in buff_amd/ imports it.

Segment<T> (database/src/segment.rs:36-44) is written by bincode 1.x
(`bincode = "1.1.3"`, database/Cargo.toml:34; Segment::convert_to_bytes /
get_byte_size / convert_from_bytes, segment.rs:84-101). bincode is a
third-party crate absent from the reference tree; this restates its published
1.x format with the default options of `bincode::serialize` /
`bincode::deserialize`: little-endian fixed-width integers, u64 sequence and
string lengths, u32 enum variant indices, u8 Option tags (0 None, 1 Some),
struct fields in declaration order, trailing bytes after the value allowed.
The serde representations used (serde 1.x, also absent): SystemTime is the
struct {secs_since_epoch: u64, nanos_since_epoch: u32}; Duration is
{secs: u64, nanos: u32}; a String is its UTF-8 bytes; both time types reject
a value whose seconds plus nanos / 10^9 overflow (u64 for Duration, the
platform's i64 seconds for SystemTime on Linux).

Field order of Segment<T> (segment.rs:36-44):
    method: Option<Methods>        Methods (methods/mod.rs:40-45): variant 0
                                   Fourier, 1 Kernel(u32), 2 SparseLearning(u32),
                                   3 DeepLearning(String)
    timestamp: SystemTime
    signal: u64                    (SignalId, segment.rs:31)
    data: Vec<T>                   T = f64 (8 bytes) or f32 (4 bytes)
    time_lapse: Option<Vec<Duration>>
    prev_seg_offset: Option<Duration>

So a Segment<f64> with method, time_lapse and prev_seg_offset all None takes
1 + 12 + 8 + 8 + 8n + 1 + 1 = 31 + 8n bytes.
"""
import struct

import numpy as np

METHOD_NONE, FOURIER, KERNEL, SPARSE_LEARNING, DEEP_LEARNING = -1, 0, 1, 2, 3
I64_MAX = (1 << 63) - 1
U64_MAX = (1 << 64) - 1


class SegmentError(ValueError):
    """bincode::deserialize returned Err (convert_from_bytes maps it to Err(()))."""


def _dtype(elem_bytes):
    if elem_bytes == 8:
        return np.dtype("<f8")
    if elem_bytes == 4:
        return np.dtype("<f4")
    raise ValueError("T is f64 or f32")


def serialize(method=METHOD_NONE, dictionary_id=0, model=b"", ts_secs=0, ts_nanos=0, signal=0,
              data=(), elem_bytes=8, time_lapse=None, prev_offset=None):
    """bincode::serialize(&segment) as bytes. time_lapse: None or a list of
    (secs, nanos); prev_offset: None or (secs, nanos)."""
    out = bytearray()
    if method == METHOD_NONE:
        out += b"\x00"
    else:
        out += b"\x01" + struct.pack("<I", method)
        if method in (KERNEL, SPARSE_LEARNING):
            out += struct.pack("<I", dictionary_id)
        elif method == DEEP_LEARNING:
            out += struct.pack("<Q", len(model)) + bytes(model)
        elif method != FOURIER:
            raise ValueError("no Methods variant %d" % method)
    out += struct.pack("<QI", ts_secs, ts_nanos)
    out += struct.pack("<Q", signal)
    arr = np.ascontiguousarray(data, dtype=_dtype(elem_bytes))
    out += struct.pack("<Q", len(arr)) + arr.tobytes()
    if time_lapse is None:
        out += b"\x00"
    else:
        out += b"\x01" + struct.pack("<Q", len(time_lapse))
        for s, ns in time_lapse:
            out += struct.pack("<QI", s, ns)
    if prev_offset is None:
        out += b"\x00"
    else:
        out += b"\x01" + struct.pack("<QI", *prev_offset)
    return bytes(out)


def byte_size(**kw):
    """bincode::serialized_size (Segment::get_byte_size, segment.rs:90-95)."""
    return len(serialize(**kw))


class _Reader:
    def __init__(self, b):
        self.b, self.i = memoryview(bytes(b)), 0

    def take(self, k):
        if k > len(self.b) - self.i:
            raise SegmentError("unexpected end of input")
        v = self.b[self.i:self.i + k]
        self.i += k
        return bytes(v)

    def u8(self):
        return self.take(1)[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def option(self):
        t = self.u8()
        if t > 1:
            raise SegmentError("invalid Option tag %d" % t)
        return t == 1

    def duration(self, secs_limit):
        s, ns = self.u64(), self.u32()
        if s + ns // 1_000_000_000 > secs_limit:
            raise SegmentError("overflow deserializing time")
        return s, ns


def parse(b, elem_bytes=8):
    """bincode::deserialize::<Segment<T>>(b) as a dict; raises SegmentError."""
    r = _Reader(b)
    seg = {"method": METHOD_NONE, "dictionary_id": 0, "model": b""}
    if r.option():
        v = r.u32()
        if v > DEEP_LEARNING:
            raise SegmentError("invalid Methods variant %d" % v)
        seg["method"] = v
        if v in (KERNEL, SPARSE_LEARNING):
            seg["dictionary_id"] = r.u32()
        elif v == DEEP_LEARNING:
            n = r.u64()
            model = r.take(n)
            try:
                model.decode("utf-8")
            except UnicodeDecodeError as e:
                raise SegmentError("invalid UTF-8") from e
            seg["model"] = model
    seg["ts_secs"], seg["ts_nanos"] = r.duration(I64_MAX)
    seg["signal"] = r.u64()
    n = r.u64()
    dt = _dtype(elem_bytes)
    if n > (len(r.b) - r.i) // dt.itemsize:
        raise SegmentError("unexpected end of input")
    seg["data"] = np.frombuffer(r.take(n * dt.itemsize), dtype=dt).copy()
    seg["time_lapse"] = None
    if r.option():
        k = r.u64()
        seg["time_lapse"] = [r.duration(U64_MAX) for _ in range(k)]
    seg["prev_offset"] = r.duration(U64_MAX) if r.option() else None
    seg["consumed"] = r.i
    return seg
