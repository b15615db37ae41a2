"""`BuffSliceCompress` with the reference's method names, over the C ABI.

Mirrors database/src/compress/buff_slice.rs:222-232 (struct + constructor,
the same {chunksize, batchsize, scale} fields as SplitBDDoubleCompress) and
the methods its two drivers call (buff_slice.rs:30-171). Streams are numpy
uint8 arrays; every method runs on the GPU through a `buff_sd_buff_slice_*`
entry point of libbuffgpu.so. Where the reference only prints a count, a sum
or a max, the method returns it (filters also return the bitmap in
BitVec<u32> word layout, maxima the argmax count).
"""
import ctypes

from . import _native as N
from .split_double import SplitBDDoubleCompress, _u8


class BuffSliceCompress(SplitBDDoubleCompress):
    """Shares the stream helpers of SplitBDDoubleCompress; only the
    buff_slice_* methods below belong to this struct in the reference."""

    def buff_slice_encode(self, data):  # buff_slice.rs:298
        return self._encode("buff_sd_buff_slice_encode", data)

    def buff_slice_decode(self, b):  # buff_slice.rs:439
        return self._decode("buff_sd_buff_slice_decode", b)

    def buff_slice_decode_bs(self, b):  # buff_slice.rs:1184
        return self._decode("buff_sd_buff_slice_decode_bs", b)

    def buff_slice_range_filter(self, b, pred):  # buff_slice.rs:1348
        return self._filter("buff_sd_buff_slice_range_filter", b, pred)

    def buff_slice_equal_filter(self, b, pred):  # buff_slice.rs:1665
        return self._filter("buff_sd_buff_slice_equal_filter", b, pred)

    def buff_slice_range_filter_nosimd(self, b, pred):  # buff_slice.rs:1753
        return self._filter("buff_sd_buff_slice_range_filter_nosimd", b, pred)

    def buff_slice_equal_filter_nosimd(self, b, pred):  # buff_slice.rs:1873
        return self._filter("buff_sd_buff_slice_equal_filter_nosimd", b, pred)

    def _sum(self, fn, b):
        b = _u8(b)
        s = ctypes.c_double()
        N.check(getattr(N.lib(), fn)(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b), ctypes.byref(s)))
        return s.value

    def buff_slice_buffsum(self, b):  # buff_slice.rs:702
        return self._sum("buff_sd_buff_slice_buffsum", b)

    def buff_slice_sum(self, b):  # buff_slice.rs:1101
        return self._sum("buff_sd_buff_slice_sum", b)

    def _max(self, fn, b):
        b = _u8(b)
        m, c = ctypes.c_double(), ctypes.c_uint64()
        N.check(getattr(N.lib(), fn)(ctypes.byref(self._sd), N.ptr(b, ctypes.c_uint8), len(b), ctypes.byref(m),
                                     ctypes.byref(c)))
        return m.value, c.value

    def buff_slice_buffmax(self, b):  # buff_slice.rs:548: (max value, argmax count)
        return self._max("buff_sd_buff_slice_buffmax", b)

    def buff_slice_max(self, b):  # buff_slice.rs:816: (max value, argmax count)
        return self._max("buff_sd_buff_slice_max", b)


__all__ = ["BuffSliceCompress"]
