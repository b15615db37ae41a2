"""`SplitBDDoubleCompress` with the reference's method names, over the C ABI.

Mirrors database/src/compress/split_double.rs:29-39 (struct + constructor) and
the methods the BUFF drivers call (buff_simd.rs:167-243, 28-95). Byte streams
are numpy uint8 arrays (the reference's `Vec<u8>`); every method runs on the
GPU through a `buff_sd_*` entry point of libbuffgpu.so. Where the reference
only prints a filter count or a max, the method returns it (filters also
return the bitmap in BitVec<u32> word layout).
"""
import ctypes

import numpy as np

from . import _native as N


def _u8(b):
    return np.ascontiguousarray(b, dtype=np.uint8)


def _stream_n(b):
    return int.from_bytes(bytes(b[8:12]), "little")


class SplitBDDoubleCompress:
    def __init__(self, chunksize, batchsize, scale):
        # split_double.rs:37
        self.chunksize = chunksize
        self.batchsize = batchsize
        self.scale = scale
        self._sd = N.SplitDouble(chunksize, batchsize, scale)

    def _encode(self, fn, data):
        x = np.ascontiguousarray(data, dtype=np.float64)
        L = N.lib()
        n = ctypes.c_uint64()
        N.check(getattr(L, fn)(ctypes.byref(self._sd), N.ptr(x, ctypes.c_double), len(x), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint8)
        N.check(getattr(L, fn)(ctypes.byref(self._sd), N.ptr(x, ctypes.c_double), len(x), N.ptr(out, ctypes.c_uint8),
                               len(out), ctypes.byref(n)))
        return out

    def byte_fixed_encode(self, data):  # split_double.rs:298
        return self._encode("buff_sd_byte_fixed_encode", data)

    def buff_simd256_encode(self, data):  # buff_simd.rs:3116
        return self._encode("buff_sd_buff_simd256_encode", data)

    def _decode(self, fn, b, *extra):
        b = _u8(b)
        n = _stream_n(b)
        out = np.zeros(max(n, 1), dtype=np.float64)
        N.check(getattr(N.lib(), fn)(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b), *extra,
                                     N.ptr(out, ctypes.c_double), len(out)))
        return out[:n]

    def byte_fixed_decode(self, b):  # split_double.rs:1243
        return self._decode("buff_sd_byte_fixed_decode", b)

    def buff_simd256_decode(self, b):  # buff_simd.rs:3252
        return self._decode("buff_sd_buff_simd256_decode", b)

    def byte_residue_encode(self, x):  # split_double.rs:624
        return self._encode("buff_sd_byte_residue_encode", x)

    def _residue_decode(self, fn, b, *extra):
        b = _u8(b)
        out = np.zeros(max(_stream_n(b), 1), dtype=np.float64)
        got = ctypes.c_uint64()
        N.check(getattr(N.lib(), fn)(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b), *extra,
                                     N.ptr(out, ctypes.c_double), len(out), ctypes.byref(got)))
        return out[: got.value]

    def byte_residue_decode(self, b):  # split_double.rs:1353
        return self._residue_decode("buff_sd_byte_residue_decode", b)

    def byte_residue_decode_with_precision(self, b, precision):  # split_double.rs:1612
        return self._residue_decode("buff_sd_byte_residue_decode_with_precision", b, precision)

    def byte_residue_sum(self, b):  # split_double.rs:2210
        b = _u8(b)
        s = ctypes.c_double()
        N.check(N.lib().buff_sd_byte_residue_sum(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                                 ctypes.byref(s)))
        return s.value

    def _filter(self, fn, b, pred):
        b = _u8(b)
        n = _stream_n(b)
        words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        cnt = ctypes.c_uint64()
        N.check(getattr(N.lib(), fn)(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b), float(pred),
                                     N.ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt)))
        return words[: (n + 31) // 32], cnt.value

    def buff_range_filter(self, b, pred):  # buff_simd.rs:4507
        return self._filter("buff_sd_buff_range_filter", b, pred)

    def buff_equal_filter(self, b, pred):  # buff_simd.rs:4925
        return self._filter("buff_sd_buff_equal_filter", b, pred)

    def buff_simd_range_filter(self, b, pred):  # buff_simd.rs:3361
        return self._filter("buff_sd_buff_simd_range_filter", b, pred)

    def buff_simd_equal_filter(self, b, pred):  # buff_simd.rs:3549
        return self._filter("buff_sd_buff_simd_equal_filter", b, pred)

    def byte_fixed_sum(self, b):  # split_double.rs:2127
        b = _u8(b)
        s = ctypes.c_double()
        N.check(N.lib().buff_sd_byte_fixed_sum(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                               ctypes.byref(s)))
        return s.value

    def byte_residue_sum_with_precision(self, b, precision):  # split_double.rs:2409
        b = _u8(b)
        s = ctypes.c_double()
        N.check(N.lib().buff_sd_byte_residue_sum_with_precision(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8),
                                                                len(b), int(precision), ctypes.byref(s)))
        return s.value

    def byte_fixed_max(self, b):  # split_double.rs:3919
        b = _u8(b)
        m = ctypes.c_double()
        N.check(N.lib().buff_sd_byte_fixed_max(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                               ctypes.byref(m)))
        return m.value

    # ---- buff-major (buff_simd.rs:317-2861) -------------------------------------

    def buff_encode_majority(self, data):  # buff_simd.rs:317
        return self._encode("buff_sd_buff_encode_majority", data)

    def buff_decode_majority(self, b):  # buff_simd.rs:633 (empty when the stream ends in sparse columns)
        b = _u8(b)
        n = _stream_n(b)
        out = np.zeros(max(n, 1), dtype=np.float64)
        got = ctypes.c_uint64()
        N.check(N.lib().buff_sd_buff_decode_majority(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                                     N.ptr(out, ctypes.c_double), len(out), ctypes.byref(got)))
        return out[: got.value]

    def buff_range_filter_majority(self, b, pred):  # buff_simd.rs:863
        return self._filter("buff_sd_buff_range_filter_majority", b, pred)

    def buff_equal_filter_majority(self, b, pred):  # buff_simd.rs:1916
        return self._filter("buff_sd_buff_equal_filter_majority", b, pred)

    def buff_sum_majority(self, b):  # buff_simd.rs:2294
        b = _u8(b)
        s = ctypes.c_double()
        N.check(N.lib().buff_sd_buff_sum_majority(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                                  ctypes.byref(s)))
        return s.value

    def buff_max_majority(self, b):  # buff_simd.rs:2444
        b = _u8(b)
        m = ctypes.c_double()
        N.check(N.lib().buff_sd_buff_max_majority(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b),
                                                  ctypes.byref(m)))
        return m.value

    # ---- buff-simd "_with_slice" filters (buff_simd.rs:3736-4025) -----------------

    def buff_simd_range_filter_with_slice(self, b, pred):  # buff_simd.rs:3736
        return self._filter("buff_sd_buff_simd_range_filter_with_slice", b, pred)

    def buff_simd_equal_filter_with_slice(self, b, pred):  # buff_simd.rs:3888
        return self._filter("buff_sd_buff_simd_equal_filter_with_slice", b, pred)
