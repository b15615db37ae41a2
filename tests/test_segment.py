"""The Segment container (segment.rs:36-44) as bincode 1.x writes it, on the CPU.

The oracle (oracle/refsegment.py) is pinned by byte strings written out by
hand from the format (bincode 1.x default options; serde's SystemTime and
Duration structs) and by the size the reference's MiB/s line divides by,
31 + 8n for a Segment<f64> whose Options are None. The library's serializer and
parser (host code of libbuffgpu, called through the C ABI) must match the
oracle byte for byte and field for field, and reject exactly the inputs the
oracle rejects (early end at every cut, Option tags above 1, Methods variants
above 3, non-UTF-8 model names, time overflow), with trailing bytes allowed.
"""
import struct

import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd.segment import DEEP_LEARNING, FOURIER, KERNEL, METHOD_NONE, SPARSE_LEARNING, Segment
from oracle import refsegment as R


def _hex(s):
    return bytes.fromhex(s.replace(" ", ""))


# hand-written from the format: Option tag | SystemTime (u64 secs, u32 nanos) | u64 signal |
# u64 len + f64 LE | Option tag | Option tag
GOLDEN = [
    (dict(ts_secs=1, ts_nanos=2, signal=3, data=[1.5]),
     "00" "0100000000000000" "02000000" "0300000000000000" "0100000000000000" "000000000000F83F" "00" "00"),
    (dict(method=KERNEL, dictionary_id=7, data=[]),
     "01" "01000000" "07000000" "0000000000000000" "00000000" "0000000000000000" "0000000000000000" "00" "00"),
    (dict(method=DEEP_LEARNING, model=b"ab", signal=0x0102030405060708, data=[-2.0],
          time_lapse=[(5, 6)], prev_offset=(9, 10)),
     "01" "03000000" "0200000000000000" "6162" "0000000000000000" "00000000" "0807060504030201"
     "0100000000000000" "00000000000000C0" "01" "0100000000000000" "0500000000000000" "06000000"
     "01" "0900000000000000" "0A000000"),
    (dict(method=FOURIER, ts_secs=0xFFFFFFFF, data=[0.25], elem_bytes=4),
     "01" "00000000" "FFFFFFFF00000000" "00000000" "0000000000000000" "0100000000000000" "0000803E" "00" "00"),
]


@pytest.mark.parametrize("i", range(len(GOLDEN)))
def test_oracle_golden_bytes(i):
    kw, hx = GOLDEN[i]
    assert R.serialize(**kw) == _hex(hx)
    got = R.parse(_hex(hx), kw.get("elem_bytes", 8))
    assert list(got["data"]) == list(kw.get("data", []))
    assert got["consumed"] == len(_hex(hx))


def test_oracle_size_is_31_plus_8n():
    for n in (0, 1, 7, 1000):
        assert R.byte_size(data=np.zeros(n)) == 31 + 8 * n
        assert R.byte_size(data=np.zeros(n), elem_bytes=4) == 31 + 4 * n


def _seg(kw):
    """Segment (the library's mirror) from oracle keyword arguments."""
    return Segment(data=np.asarray(kw.get("data", []), dtype=np.float64 if kw.get("elem_bytes", 8) == 8
                                   else np.float32),
                   signal=kw.get("signal", 0), timestamp=(kw.get("ts_secs", 0), kw.get("ts_nanos", 0)),
                   method=kw.get("method", METHOD_NONE), dictionary_id=kw.get("dictionary_id", 0),
                   model=kw.get("model", b""), time_lapse=kw.get("time_lapse"), prev_seg_offset=kw.get("prev_offset"),
                   elem_bytes=kw.get("elem_bytes", 8))


def _random_kw(rng):
    eb = int(rng.choice([8, 4]))
    m = int(rng.integers(-1, 4))
    kw = dict(method=m, ts_secs=int(rng.integers(0, 1 << 62)), ts_nanos=int(rng.integers(0, 10 ** 9)),
              signal=int(rng.integers(0, 1 << 63)), elem_bytes=eb,
              data=rng.normal(size=int(rng.integers(0, 50))) * 10 ** rng.uniform(-3, 6))
    if m in (KERNEL, SPARSE_LEARNING):
        kw["dictionary_id"] = int(rng.integers(0, 1 << 32))
    if m == DEEP_LEARNING:
        kw["model"] = "model-%d-é" % int(rng.integers(0, 1000))
        kw["model"] = kw["model"].encode()
    if rng.random() < 0.5:
        kw["time_lapse"] = [(int(rng.integers(0, 1 << 40)), int(rng.integers(0, 10 ** 9)))
                            for _ in range(int(rng.integers(0, 5)))]
    if rng.random() < 0.5:
        kw["prev_offset"] = (int(rng.integers(0, 1 << 40)), int(rng.integers(0, 10 ** 9)))
    return kw


@pytest.mark.parametrize("i", range(len(GOLDEN)))
def test_library_golden_bytes(i):
    kw, hx = GOLDEN[i]
    s = _seg(kw)
    assert s.to_bytes() == _hex(hx)
    assert s.byte_size() == len(_hex(hx))


def test_library_matches_oracle_random():
    rng = np.random.default_rng(0)
    for _ in range(200):
        kw = _random_kw(rng)
        ref = R.serialize(**kw)
        s = _seg(kw)
        assert s.to_bytes() == ref
        assert s.byte_size() == len(ref)
        got = N.segment_parse(ref + b"\x07\x07", kw["elem_bytes"])   # trailing bytes are allowed
        want = R.parse(ref + b"\x07\x07", kw["elem_bytes"])
        for k in ("method", "dictionary_id", "model", "ts_secs", "ts_nanos", "signal", "time_lapse", "prev_offset",
                  "consumed"):
            assert got[k] == want[k], k
        assert np.array_equal(got["data"].view(np.uint8), want["data"].view(np.uint8))
        back = Segment.from_bytes(ref, kw["elem_bytes"])
        assert back.to_bytes() == ref


def _both_reject(b, eb=8):
    with pytest.raises(R.SegmentError):
        R.parse(b, eb)
    with pytest.raises(N.BuffError) as e:
        N.segment_parse(b, eb)
    assert e.value.code == N.BUFF_E_INVALID


def test_rejections_match_oracle():
    kw, hx = GOLDEN[2]
    full = _hex(hx)
    for cut in range(len(full)):   # every early end
        _both_reject(full[:cut])
    _both_reject(b"\x02" + full[1:])                                          # Option tag 2
    _both_reject(b"\x01" + struct.pack("<I", 4) + full[5:])                    # Methods variant 4
    bad_utf8 = R.serialize(method=DEEP_LEARNING, model=b"\xff\xfe", data=[1.0])
    _both_reject(bad_utf8)
    for secs, nanos in ((1 << 63, 0), ((1 << 63) - 1, 10 ** 9)):               # SystemTime past i64 seconds
        _both_reject(R.serialize(ts_secs=secs, ts_nanos=nanos, data=[]))
    ok = R.serialize(ts_secs=(1 << 63) - 1, ts_nanos=10 ** 9 - 1, data=[])
    assert N.segment_parse(ok)["ts_secs"] == (1 << 63) - 1
    _both_reject(R.serialize(data=[], prev_offset=((1 << 64) - 1, 10 ** 9)))   # Duration past u64 seconds
    _both_reject(R.serialize(data=[], time_lapse=[(1, 2), ((1 << 64) - 1, 2 * 10 ** 9)]))
    big = bytearray(R.serialize(data=[1.0, 2.0]))
    big[21:29] = struct.pack("<Q", 1 << 60)                                    # data length past the input
    _both_reject(bytes(big))


def test_oracle_rejects_what_library_rejects_on_random_bytes():
    rng = np.random.default_rng(3)
    for _ in range(300):
        b = bytes(rng.integers(0, 256, int(rng.integers(0, 60)), dtype=np.uint8))
        try:
            want = R.parse(b)
        except R.SegmentError:
            with pytest.raises(N.BuffError):
                N.segment_parse(b)
            continue
        got = N.segment_parse(b)
        assert got["consumed"] == want["consumed"]
