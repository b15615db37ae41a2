"""byte_fixed_encode on a serialized Segment<T> (buff_segment_encode) on the GPU.

The reference's encoders take a &mut Segment<T> and read its data
(split_double.rs:298-318). A Segment serialized by the oracle
(oracle/refsegment.py, bincode 1.x) with f64 or f32 data, methods and time
fields set, and trailing bytes after it, is encoded through the C ABI; the
chunk's stream must equal the oracle's encode of the same values as float64
(f32 -> f64 is exact), in the plain and the buff-major layout.
"""
import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg
from oracle import refsegment as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("elem_bytes", [8, 4])
def test_encode_segment_matches_oracle(gpu_ctx, oracle, elem_bytes):
    x = dg.random_walk(100_003, 3, 21)
    if elem_bytes == 4:
        x = x.astype(np.float32)
    b = R.serialize(method=R.KERNEL, dictionary_id=9, ts_secs=1_700_000_000, ts_nanos=5, signal=42, data=x,
                    elem_bytes=elem_bytes, time_lapse=[(1, 0), (2, 500)], prev_offset=(60, 0)) + b"\x00\x01"
    print("segment", elem_bytes, len(b), flush=True)
    ch = gpu_ctx.encode_segment(b, 1000, elem_bytes=elem_bytes)
    try:
        assert np.array_equal(ch.serialize(), oracle.encode(np.asarray(x, dtype=np.float64), 1000))
    finally:
        ch.free()


def test_encode_segment_majority(gpu_ctx, oracle):
    x = dg.outlier_heavy(70_000, 3, 22)
    b = R.serialize(data=x)
    ch = gpu_ctx.encode_segment(b, 1000, flags=N.BUFF_MAJORITY)
    try:
        assert np.array_equal(ch.serialize(), oracle.major_encode(x, 1000))
    finally:
        ch.free()


def test_encode_segment_rejects_bad_bytes(gpu_ctx):
    b = R.serialize(data=[1.0, 2.0])
    with pytest.raises(N.BuffError) as e:
        gpu_ctx.encode_segment(b[:-3], 10)
    assert e.value.code == N.BUFF_E_INVALID
