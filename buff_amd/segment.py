"""Segment<T> (database/src/segment.rs:36-44) and its bincode form, over the C ABI.

Segment.to_bytes / byte_size / from_bytes are convert_to_bytes, get_byte_size
and convert_from_bytes (segment.rs:84-101); the bytes are written and parsed by
libbuffgpu (buff_segment_serialize / buff_segment_parse). Methods variants
(methods/mod.rs:40-45): METHOD_NONE for None, else FOURIER, KERNEL(dictionary
id), SPARSE_LEARNING(dictionary id), DEEP_LEARNING(model name).
"""
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _native as N

METHOD_NONE, FOURIER, KERNEL, SPARSE_LEARNING, DEEP_LEARNING = -1, 0, 1, 2, 3


@dataclass
class Segment:
    data: np.ndarray
    signal: int = 0
    timestamp: Tuple[int, int] = (0, 0)           # SystemTime since UNIX_EPOCH: (secs, nanos)
    method: int = METHOD_NONE
    dictionary_id: int = 0                        # Kernel / SparseLearning
    model: bytes = b""                            # DeepLearning
    time_lapse: Optional[List[Tuple[int, int]]] = None
    prev_seg_offset: Optional[Tuple[int, int]] = None
    elem_bytes: int = 8                           # T = f64 (8) or f32 (4)

    def to_bytes(self) -> bytes:
        return N.segment_serialize(self)

    def byte_size(self) -> int:
        return N.segment_byte_size(self)

    @classmethod
    def from_bytes(cls, b, elem_bytes=8) -> "Segment":
        d = N.segment_parse(b, elem_bytes)
        return cls(data=d["data"], signal=d["signal"], timestamp=(d["ts_secs"], d["ts_nanos"]), method=d["method"],
                   dictionary_id=d["dictionary_id"], model=d["model"], time_lapse=d["time_lapse"],
                   prev_seg_offset=d["prev_offset"], elem_bytes=elem_bytes)
