"""ctypes loader for the oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg. The product (buff_amd/) never
imports it.

Each wrapper returns numpy arrays / Python scalars and raises RefPanic where
the reference would panic (status 4) and RefUnsupported where the stream shape
is outside the cited branch (status 5).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BUFF_ORACLE_LIB: load another build of refcpu.cpp (the planted-mutation test
# points it at mutated copies); nothing is built then.
_LIB_OVERRIDE = os.environ.get("BUFF_ORACLE_LIB")
_LIB_PATH = _LIB_OVERRIDE or os.path.join(_HERE, "librefcpu.so")

REF_OK, REF_PANIC, REF_UNSUPPORTED, REF_SMALL_BUFFER = 0, 4, 5, 6


class RefPanic(RuntimeError):
    """The reference would panic on this input."""


class RefUnsupported(RuntimeError):
    """The cited reference branch does not cover this stream."""


def build(force=False):
    """Compile librefcpu.so from refcpu.cpp with the committed Makefile (make
    rebuilds it when a source is newer)."""
    if _LIB_OVERRIDE:
        return
    subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        u8p, u32p, u64p, i64p, f64p = (P(ctypes.c_uint8), P(ctypes.c_uint32), P(ctypes.c_uint64),
                                       P(ctypes.c_int64), P(ctypes.c_double))
        sig = {
            "ref_prec_of_scale": (ctypes.c_int, [ctypes.c_uint64]),
            "ref_get_precision_bound": (ctypes.c_double, [ctypes.c_int32]),
            "ref_precision_exp": (ctypes.c_int32, [ctypes.c_double]),
            "ref_precision_map": (ctypes.c_int, [ctypes.c_int32, P(ctypes.c_int32)]),
            "ref_fetch_fixed_aligned": (ctypes.c_int64, [ctypes.c_double, ctypes.c_int32, ctypes.c_uint64]),
            "ref_powi": (ctypes.c_double, [ctypes.c_double, ctypes.c_int32]),
            "ref_bitpack_write": (ctypes.c_int, [u32p, u32p, u8p, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p]),
            "ref_bitpack_read": (ctypes.c_int, [u8p, ctypes.c_uint64, u32p, ctypes.c_uint64, u32p]),
            "ref_encode": (ctypes.c_int, [f64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, u8p,
                                          ctypes.c_uint64, u64p, i64p]),
            "ref_decode": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, f64p,
                                          ctypes.c_uint64, u64p]),
            "ref_filter": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                          u32p, ctypes.c_uint64, u64p, u64p]),
            "ref_filter_cond": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, u64p,
                                               ctypes.c_uint64, ctypes.c_int, u32p, ctypes.c_uint64, u64p]),
            "ref_decode_cond": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, u64p, ctypes.c_uint64,
                                               ctypes.c_int, f64p, ctypes.c_uint64, u64p]),
            "ref_max_range": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_int, u64p, f64p, ctypes.c_uint64, u64p, u32p, ctypes.c_uint64]),
            "ref_simd_filter": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                               ctypes.c_double, u32p, ctypes.c_uint64, u64p]),
            "ref_sum": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, f64p, u64p]),
            "ref_max": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, f64p, u64p, u64p]),
            "ref_decode_with_precision": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                         f64p, ctypes.c_uint64, u64p]),
            "ref_sum_with_precision": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, f64p,
                                                      u64p]),
            "ref_major_encode": (ctypes.c_int, [f64p, ctypes.c_uint64, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p,
                                                i64p]),
            "ref_major_decode": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, f64p, ctypes.c_uint64, u64p]),
            "ref_major_filter": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                                u32p, ctypes.c_uint64, u64p]),
            "ref_major_sum": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, f64p, u64p]),
            "ref_major_max": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, f64p, u64p, u64p]),
            "ref_slice_encode": (ctypes.c_int, [f64p, ctypes.c_uint64, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p]),
            "ref_slice_decode": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_int, u64p, ctypes.c_uint64, f64p,
                                                ctypes.c_uint64, u64p]),
            "ref_slice_sum": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_int, f64p]),
            "ref_slice_max": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_int, f64p, u64p, u32p, ctypes.c_uint64]),
            "ref_slice_max_range": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                   ctypes.c_uint32, f64p, ctypes.c_uint64, u64p, u32p,
                                                   ctypes.c_uint64]),
            "ref_slice_filter": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                                u64p, ctypes.c_uint64, u32p, ctypes.c_uint64, u64p]),
            "ref_snappy_compress": (ctypes.c_int, [u8p, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p]),
            "ref_snappy_decompress": (ctypes.c_int, [u8p, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p]),
            "ref_residue_encode": (ctypes.c_int, [f64p, ctypes.c_uint64, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p,
                                                  u32p]),
            "ref_residue_decode": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_int32, f64p, ctypes.c_uint64, u64p]),
            "ref_residue_sum": (ctypes.c_int, [u8p, ctypes.c_uint64, ctypes.c_int32, f64p, u64p, u32p]),
            "ref_roaring_serialize": (ctypes.c_int, [u32p, ctypes.c_uint64, u8p, ctypes.c_uint64, u64p]),
            "ref_roaring_deserialize": (ctypes.c_int, [u8p, ctypes.c_uint64, u32p, ctypes.c_uint64, u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _check(rc):
    if rc == REF_PANIC:
        raise RefPanic("reference panics on this input")
    if rc == REF_UNSUPPORTED:
        raise RefUnsupported("outside the cited reference branch")
    if rc != REF_OK:
        raise RuntimeError("oracle status %d" % rc)


def prec_of_scale(scale):
    return lib().ref_prec_of_scale(scale)


def get_precision_bound(p):
    return lib().ref_get_precision_bound(p)


def precision_exp(precision):
    return lib().ref_precision_exp(precision)


def precision_map(p):
    d = ctypes.c_int32()
    _check(lib().ref_precision_map(p, ctypes.byref(d)))
    return d.value


def fetch_fixed_aligned(x, precision_exp_, dlen):
    return lib().ref_fetch_fixed_aligned(float(x), precision_exp_, dlen)


def powi(a, b):
    return lib().ref_powi(a, b)


def bitpack_write(ops):
    """ops: list of (kind, value, bits); kind in {'write','write_bits','write_byte','finish_write_byte'}."""
    kinds = {"write": 0, "write_bits": 1, "write_byte": 2, "finish_write_byte": 3}
    k = len(ops)
    vals = np.array([o[1] for o in ops], dtype=np.uint32)
    bits = np.array([o[2] for o in ops], dtype=np.uint32)
    kind = np.array([kinds[o[0]] for o in ops], dtype=np.uint8)
    cap = 64 + sum(int(b) for b in bits) // 8 + k
    out = np.zeros(cap, dtype=np.uint8)
    n = ctypes.c_uint64()
    _check(lib().ref_bitpack_write(_ptr(vals, ctypes.c_uint32), _ptr(bits, ctypes.c_uint32),
                                   _ptr(kind, ctypes.c_uint8), k, _ptr(out, ctypes.c_uint8), cap, ctypes.byref(n)))
    return out[: n.value].copy()


def bitpack_read(buf, bits):
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    vals = np.zeros(len(bits), dtype=np.uint32)
    _check(lib().ref_bitpack_read(_ptr(buf, ctypes.c_uint8), len(buf), _ptr(bits, ctypes.c_uint32), len(bits),
                                  _ptr(vals, ctypes.c_uint32)))
    return vals


def encode(x, scale, flip=False, with_minmax=False):
    """byte_fixed_encode (flip=False) / buff_simd256_encode (flip=True) -> uint8 stream."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = ctypes.c_uint64()
    mm = np.zeros(2, dtype=np.int64)
    L = lib()
    _check(L.ref_encode(_ptr(x, ctypes.c_double), len(x), scale, int(flip), None, 0, ctypes.byref(n),
                        _ptr(mm, ctypes.c_int64)))
    out = np.zeros(n.value, dtype=np.uint8)
    _check(L.ref_encode(_ptr(x, ctypes.c_double), len(x), scale, int(flip), _ptr(out, ctypes.c_uint8), n.value,
                        ctypes.byref(n), _ptr(mm, ctypes.c_int64)))
    return (out, int(mm[0]), int(mm[1])) if with_minmax else out


def header(stream):
    s = np.asarray(stream, dtype=np.uint8)
    h = s[:20].tobytes()
    base = int.from_bytes(h[0:8], "little", signed=True)
    n = int.from_bytes(h[8:12], "little")
    ilen = int.from_bytes(h[12:16], "little")
    dlen = int.from_bytes(h[16:20], "little")
    return base, n, ilen, dlen


def decode(stream, scale, flip=False):
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    out = np.zeros(max(n, 1), dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_decode(_ptr(s, ctypes.c_uint8), len(s), scale, int(flip), _ptr(out, ctypes.c_double),
                            len(out), ctypes.byref(got)))
    return out[: got.value]


def buff_filter(stream, scale, op, pred):
    """op 'range' (x > pred) or 'equal'. Returns (u32 words, count, candidates[9])."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    cnt = ctypes.c_uint64()
    cand = np.zeros(9, dtype=np.uint64)
    _check(lib().ref_filter(_ptr(s, ctypes.c_uint8), len(s), scale, 0 if op == "range" else 1, float(pred),
                            _ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt), _ptr(cand, ctypes.c_uint64)))
    return words[: (n + 31) // 32], cnt.value, cand


def filter_cond(stream, scale, pred, rows, major=False):
    """buff_range_filter_condition (major: buff_range_filter_majority_condition) over ascending
    row ids. Returns (u32 words, count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    cnt = ctypes.c_uint64()
    _check(lib().ref_filter_cond(_ptr(s, ctypes.c_uint8), len(s), scale, float(pred), _ptr(r, ctypes.c_uint64), len(r),
                                 int(major), _ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt)))
    return words[: (n + 31) // 32], cnt.value


def decode_cond(stream, scale, rows, major=False):
    """buff_decode_condition (major: buff_major_decode_condition) over ascending row ids: the
    rows' values in order."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    out = np.zeros(max(len(r), 1), dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_decode_cond(_ptr(s, ctypes.c_uint8), len(s), scale, _ptr(r, ctypes.c_uint64), len(r),
                                 int(major), _ptr(out, ctypes.c_double), len(out), ctypes.byref(got)))
    return out[: got.value]


def max_range(stream, s, e, window, major=False):
    """buff_max_range (major: buff_max_majority_range) over rows [s, e) in windows of `window` rows:
    (window max offsets, their f64 values, argmax u32 words)."""
    st = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(st)[1]
    cap = max((e - s) // max(window, 1) + 2, 1) if e >= s else 1
    offs = np.zeros(cap, dtype=np.uint64)
    vals = np.zeros(cap, dtype=np.float64)
    nwin = ctypes.c_uint64()
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    _check(lib().ref_max_range(_ptr(st, ctypes.c_uint8), len(st), s, e, window, int(major), _ptr(offs, ctypes.c_uint64),
                               _ptr(vals, ctypes.c_double), cap, ctypes.byref(nwin), _ptr(words, ctypes.c_uint32),
                               len(words)))
    k = nwin.value
    return offs[:k], vals[:k], words[: (n + 31) // 32]


def simd_filter(stream, scale, op, pred):
    """AVX2 buff-simd filter over a flipped stream. Returns (u32 words, count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    cnt = ctypes.c_uint64()
    _check(lib().ref_simd_filter(_ptr(s, ctypes.c_uint8), len(s), scale, 0 if op == "range" else 1, float(pred),
                                 _ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt)))
    return words[: (n + 31) // 32], cnt.value


def byte_fixed_sum(stream, scale):
    """byte_fixed_sum -> (f64 sum, per-sub-column u64 sums[9])."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    ps = np.zeros(9, dtype=np.uint64)
    _check(lib().ref_sum(_ptr(s, ctypes.c_uint8), len(s), scale, ctypes.byref(out), _ptr(ps, ctypes.c_uint64)))
    return out.value, ps


def byte_fixed_max(stream, scale):
    """byte_fixed_max -> (printed max value, max stored offset, argmax count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    off = ctypes.c_uint64()
    cnt = ctypes.c_uint64()
    _check(lib().ref_max(_ptr(s, ctypes.c_uint8), len(s), scale, ctypes.byref(out), ctypes.byref(off),
                         ctypes.byref(cnt)))
    return out.value, off.value, cnt.value


def decode_with_precision(stream, scale, precision):
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    out = np.zeros(max(n, 1), dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_decode_with_precision(_ptr(s, ctypes.c_uint8), len(s), scale, precision,
                                           _ptr(out, ctypes.c_double), len(out), ctypes.byref(got)))
    return out[: got.value]


def sum_with_precision(stream, scale, precision):
    """byte_residue_sum_with_precision -> (f64 sum, u64 sums of the sub-columns read and of the
    R % 8-bit fields, index by read order)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    ps = np.zeros(9, dtype=np.uint64)
    _check(lib().ref_sum_with_precision(_ptr(s, ctypes.c_uint8), len(s), scale, precision, ctypes.byref(out),
                                        _ptr(ps, ctypes.c_uint64)))
    return out.value, ps


# --- RAPG residue format (split_double.rs:624-954, 1353-1893, 2210-2629) ------

def residue_encode(x, scale):
    """byte_residue_encode -> (uint8 stream, outlier columns)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = ctypes.c_uint64()
    cols = ctypes.c_uint32()
    L = lib()
    _check(L.ref_residue_encode(_ptr(x, ctypes.c_double), len(x), scale, None, 0, ctypes.byref(n), ctypes.byref(cols)))
    out = np.zeros(max(n.value, 1), dtype=np.uint8)
    _check(L.ref_residue_encode(_ptr(x, ctypes.c_double), len(x), scale, _ptr(out, ctypes.c_uint8), n.value,
                                ctypes.byref(n), ctypes.byref(cols)))
    return out[: n.value], cols.value


def residue_decode(stream, precision=-1):
    """byte_residue_decode (precision < 0) / byte_residue_decode_with_precision -> float64 values."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    out = np.zeros(max(n, 1), dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_residue_decode(_ptr(s, ctypes.c_uint8), len(s), precision, _ptr(out, ctypes.c_double), len(out),
                                    ctypes.byref(got)))
    return out[: got.value]


def residue_sum(stream, precision=-1):
    """byte_residue_sum (precision < 0) / byte_residue_sum_with_precision -> (f64 sum, u64 sums of the
    columns read in order (outlier columns: their outlier bytes), outlier columns)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    ps = np.zeros(9, dtype=np.uint64)
    cols = ctypes.c_uint32()
    _check(lib().ref_residue_sum(_ptr(s, ctypes.c_uint8), len(s), precision, ctypes.byref(out),
                                 _ptr(ps, ctypes.c_uint64), ctypes.byref(cols)))
    return out.value, ps, cols.value


def roaring_serialize(values):
    """croaring Bitmap of `values` (ascending), run_optimize(), serialize() -> bytes."""
    v = np.ascontiguousarray(values, dtype=np.uint32)
    n = ctypes.c_uint64()
    L = lib()
    L.ref_roaring_serialize(_ptr(v, ctypes.c_uint32), len(v), None, 0, ctypes.byref(n))
    out = np.zeros(max(n.value, 1), dtype=np.uint8)
    _check(L.ref_roaring_serialize(_ptr(v, ctypes.c_uint32), len(v), _ptr(out, ctypes.c_uint8), n.value,
                                   ctypes.byref(n)))
    return bytes(out[: n.value])


def roaring_deserialize(b, cap=1 << 22):
    """Bitmap::deserialize -> ascending uint32 values."""
    s = np.frombuffer(bytes(b), dtype=np.uint8)
    out = np.zeros(cap, dtype=np.uint32)
    n = ctypes.c_uint64()
    _check(lib().ref_roaring_deserialize(_ptr(s, ctypes.c_uint8) if len(s) else None, len(s),
                                         _ptr(out, ctypes.c_uint32), cap, ctypes.byref(n)))
    return out[: n.value]


# --- buff-major (buff_simd.rs:317-2861) -------------------------------------

def major_encode(x, scale, with_minmax=False):
    """buff_encode_majority -> uint8 stream."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = ctypes.c_uint64()
    mm = np.zeros(2, dtype=np.int64)
    L = lib()
    _check(L.ref_major_encode(_ptr(x, ctypes.c_double), len(x), scale, None, 0, ctypes.byref(n),
                              _ptr(mm, ctypes.c_int64)))
    out = np.zeros(n.value, dtype=np.uint8)
    _check(L.ref_major_encode(_ptr(x, ctypes.c_double), len(x), scale, _ptr(out, ctypes.c_uint8), n.value,
                              ctypes.byref(n), _ptr(mm, ctypes.c_int64)))
    return (out, int(mm[0]), int(mm[1])) if with_minmax else out


def major_decode(stream, scale):
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    out = np.zeros(max(n, 1), dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_major_decode(_ptr(s, ctypes.c_uint8), len(s), scale, _ptr(out, ctypes.c_double), len(out),
                                  ctypes.byref(got)))
    return out[: got.value]


def major_filter(stream, scale, op, pred):
    """buff_range_filter_majority / buff_equal_filter_majority -> (u32 words, count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    cnt = ctypes.c_uint64()
    _check(lib().ref_major_filter(_ptr(s, ctypes.c_uint8), len(s), scale, 0 if op == "range" else 1, float(pred),
                                  _ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt)))
    return words[: (n + 31) // 32], cnt.value


def major_sum(stream, scale):
    """buff_sum_majority -> (f64 sum, exact byte-column sums[9])."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    cs = np.zeros(9, dtype=np.uint64)
    _check(lib().ref_major_sum(_ptr(s, ctypes.c_uint8), len(s), scale, ctypes.byref(out), _ptr(cs, ctypes.c_uint64)))
    return out.value, cs


def major_max(stream, scale):
    """buff_max_majority -> (printed max value, max offset, argmax count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    off = ctypes.c_uint64()
    cnt = ctypes.c_uint64()
    _check(lib().ref_major_max(_ptr(s, ctypes.c_uint8), len(s), scale, ctypes.byref(out), ctypes.byref(off),
                               ctypes.byref(cnt)))
    return out.value, off.value, cnt.value


def snappy_compress(data):
    d = np.frombuffer(bytes(data), dtype=np.uint8)
    cap = 32 + len(d) + len(d) // 6
    out = np.zeros(cap, dtype=np.uint8)
    n = ctypes.c_uint64()
    _check(lib().ref_snappy_compress(_ptr(d, ctypes.c_uint8), len(d), _ptr(out, ctypes.c_uint8), cap,
                                     ctypes.byref(n)))
    return out[: n.value].tobytes()


def snappy_decompress(data, size):
    d = np.frombuffer(bytes(data), dtype=np.uint8)
    out = np.zeros(max(size, 1), dtype=np.uint8)
    n = ctypes.c_uint64()
    _check(lib().ref_snappy_decompress(_ptr(d, ctypes.c_uint8), len(d), _ptr(out, ctypes.c_uint8), len(out),
                                       ctypes.byref(n)))
    return out[: n.value].tobytes()


# ---- buff-slice (buff_slice.rs:222-1970; buff_simd.rs:3736-4025) -------------

SLICE_KINDS = {"range": 0, "equal": 1, "range_nosimd": 2, "equal_nosimd": 3, "range_cond": 4, "smaller_cond": 5,
               "range_with_slice": 6, "equal_with_slice": 7}


def slice_encode(x, scale):
    """buff_slice_encode: the byte-slice stream."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    cap = 20 + len(x) * 8 + 8
    out = np.zeros(cap, dtype=np.uint8)
    ln = ctypes.c_uint64()
    _check(lib().ref_slice_encode(_ptr(x, ctypes.c_double), len(x), scale, _ptr(out, ctypes.c_uint8), cap,
                                  ctypes.byref(ln)))
    return out[: ln.value].copy()


def slice_decode(stream, kind="buff", rows=None):
    """buff_slice_decode (kind "buff"), buff_slice_decode_bs ("bs") or
    buff_slice_decode_condition ("cond", ascending rows)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    r = np.ascontiguousarray(rows if rows is not None else [], dtype=np.uint64)
    cap = max(len(r) if kind == "cond" else n, 1)
    out = np.zeros(cap, dtype=np.float64)
    got = ctypes.c_uint64()
    _check(lib().ref_slice_decode(_ptr(s, ctypes.c_uint8), len(s), {"buff": 0, "bs": 1, "cond": 2}[kind],
                                  _ptr(r, ctypes.c_uint64), len(r), _ptr(out, ctypes.c_double), cap,
                                  ctypes.byref(got)))
    return out[: got.value]


def slice_sum(stream, scalar=False):
    """buff_slice_buffsum (scalar: buff_slice_sum): the f64 running sum."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = ctypes.c_double()
    _check(lib().ref_slice_sum(_ptr(s, ctypes.c_uint8), len(s), int(scalar), ctypes.byref(out)))
    return out.value


def slice_max(stream, scalar=False):
    """buff_slice_buffmax (scalar: buff_slice_max): (max value, argmax count, argmax u32 words)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    mx = ctypes.c_double()
    cnt = ctypes.c_uint64()
    _check(lib().ref_slice_max(_ptr(s, ctypes.c_uint8), len(s), int(scalar), ctypes.byref(mx), ctypes.byref(cnt),
                               _ptr(words, ctypes.c_uint32), len(words)))
    return mx.value, cnt.value, words[: (n + 31) // 32]


def slice_max_range(stream, s, e, window):
    """buff_slice_max_range: (window max values, argmax u32 words)."""
    st = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(st)[1]
    cap = max((e - s) // max(window, 1) + 2, 1) if e >= s else 1
    vals = np.zeros(cap, dtype=np.float64)
    nwin = ctypes.c_uint64()
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    _check(lib().ref_slice_max_range(_ptr(st, ctypes.c_uint8), len(st), s, e, window, _ptr(vals, ctypes.c_double),
                                     cap, ctypes.byref(nwin), _ptr(words, ctypes.c_uint32), len(words)))
    return vals[: nwin.value], words[: (n + 31) // 32]


def slice_filter(stream, scale, kind, pred, rows=None):
    """One of the byte-slice filters (SLICE_KINDS; the *_with_slice kinds read a
    buff_simd256 stream): (u32 words, printed count)."""
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    n = header(s)[1]
    r = np.ascontiguousarray(rows if rows is not None else [], dtype=np.uint64)
    words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
    cnt = ctypes.c_uint64()
    _check(lib().ref_slice_filter(_ptr(s, ctypes.c_uint8), len(s), scale, float(pred), SLICE_KINDS[kind],
                                  _ptr(r, ctypes.c_uint64), len(r), _ptr(words, ctypes.c_uint32), len(words),
                                  ctypes.byref(cnt)))
    return words[: (n + 31) // 32], cnt.value
