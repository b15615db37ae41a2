"""MI355X-native BUFF (decomposed bounded floats) engine.

The hot path (encode, equality/range filters, sum, max, materialize) runs as
hand-written HIP kernels for gfx950 in libbuffgpu.so behind the C ABI of
include/buffgpu.h. This package is the Python binding of that ABI.
"""
from ._native import (  # noqa: F401
    BUFF_FLIP,
    BUFF_MAJORITY,
    BUFF_OP_EQ,
    BUFF_OP_GT,
    BUFF_OP_LT,
    BUFF_SCALAR,
    BUFF_SEMANTIC,
    BUFF_SLICE,
    BUFF_WITH_SLICE,
    BUFF_RESIDUE,
    BuffError,
    Batch,
    Chunk,
    Context,
    DeviceBuffer,
    OutOfMemory,
    RefPanic,
    Server,
    Unsupported,
    lib,
)
from .buff_slice import BuffSliceCompress  # noqa: F401
from .split_double import SplitBDDoubleCompress  # noqa: F401

__all__ = [
    "BUFF_FLIP", "BUFF_MAJORITY", "BUFF_OP_EQ", "BUFF_OP_GT", "BUFF_OP_LT", "BUFF_SCALAR", "BUFF_SEMANTIC",
    "BUFF_SLICE", "BUFF_WITH_SLICE", "BUFF_RESIDUE", "Batch", "BuffError", "Chunk", "Context", "DeviceBuffer",
    "OutOfMemory", "RefPanic", "Server", "Unsupported", "SplitBDDoubleCompress", "BuffSliceCompress", "lib",
]
