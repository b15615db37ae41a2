"""ctypes binding of libbuffgpu.so (the C ABI in include/buffgpu.h).

This is the Python side of the drop-in boundary: every call goes through the
exported `extern "C"` entry points of the HIP library. There is no CPU
fallback; loading fails loudly when the library is missing.
"""
import collections.abc
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbuffgpu.so")

BUFF_OK = 0
BUFF_E_INVALID = 1
BUFF_E_OOM = 2
BUFF_E_DEVICE = 3
BUFF_E_REF_PANIC = 4
BUFF_E_UNSUPPORTED = 5
BUFF_E_SMALL_BUFFER = 6

BUFF_FLIP = 0x1
BUFF_SEMANTIC = 0x2
BUFF_MAJORITY = 0x4
BUFF_SLICE = 0x8
BUFF_SCALAR = 0x10
BUFF_WITH_SLICE = 0x20
BUFF_RESIDUE = 0x40
BUFF_OP_GT = 0
BUFF_OP_EQ = 1
BUFF_OP_LT = 2


def construct_file_iterator_skip_newline(path, skip_val=0, delim=b","):
    """client.rs:163-182 on a file: its values (read on the host, parsed on the GPU)."""
    try:
        size = os.path.getsize(path)
    except OSError:
        size = 0   # the library reports the file it cannot open
    cap = size // 2 + 1
    out = np.zeros(max(cap, 1), dtype=np.float64)
    n = ctypes.c_uint64()
    check(lib().buff_construct_file_iterator_skip_newline(os.fsencode(path), skip_val, ord(delim),
                                                          ptr(out, ctypes.c_double), cap, ctypes.byref(n)))
    return out[: n.value]


def _opc(op):
    """Filter op name -> BUFF_OP_* ("range"/"gt", "equal"/"eq", "smaller"/"lt")."""
    if op in ("range", "gt", BUFF_OP_GT):
        return BUFF_OP_GT
    if op in ("equal", "eq", BUFF_OP_EQ):
        return BUFF_OP_EQ
    if op in ("smaller", "lt", BUFF_OP_LT):
        return BUFF_OP_LT
    raise ValueError("unknown filter op %r" % (op,))


class BuffError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("buffgpu status %d: %s" % (code, msg))
        self.code = code


class RefPanic(BuffError):
    """The reference panics on this input (BUFF_E_REF_PANIC)."""


class Unsupported(BuffError):
    """BUFF_E_UNSUPPORTED."""


class OutOfMemory(BuffError):
    """BUFF_E_OOM: the HBM budget or the device is exhausted."""


class ChunkInfo(ctypes.Structure):
    _fields_ = [
        ("base", ctypes.c_int64),
        ("n", ctypes.c_uint32),
        ("ilen", ctypes.c_uint32),
        ("dlen", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("planes", ctypes.c_uint32),
        ("tail_bits", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("scale", ctypes.c_uint64),
        ("stream_bytes", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("sparse_cols", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("decoded_rows", ctypes.c_uint64),
    ]


class Query(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("chunk", ctypes.c_uint32), ("pred", ctypes.c_double),
                ("precision", ctypes.c_int32), ("flags", ctypes.c_uint32)]


class QueryResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("count", ctypes.c_uint64),
                ("value", ctypes.c_double), ("exact", ctypes.c_double), ("max_off", ctypes.c_uint64),
                ("out_offset", ctypes.c_uint64)]


BUFF_Q_GT, BUFF_Q_EQ, BUFF_Q_SUM, BUFF_Q_MAX, BUFF_Q_MAT = 0, 1, 2, 3, 4


class TuneChoice(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("waves_per_wg", ctypes.c_uint32), ("max_wg", ctypes.c_uint32),
                ("tiles_per_wave", ctypes.c_uint32), ("prefetch", ctypes.c_uint32), ("best_us", ctypes.c_float),
                ("default_us", ctypes.c_float)]


class ServerCfg(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("slots", ctypes.c_uint32), ("arena_bytes", ctypes.c_uint64)]


class SplitDouble(ctypes.Structure):
    _fields_ = [("chunksize", ctypes.c_uint64), ("batchsize", ctypes.c_uint64), ("scale", ctypes.c_uint64)]


class SegmentDesc(ctypes.Structure):
    """buff_segment: the fields of Segment<T> (segment.rs:36-44)."""
    _fields_ = [("method", ctypes.c_int32), ("dictionary_id", ctypes.c_uint32), ("model", ctypes.c_void_p),
                ("model_len", ctypes.c_uint64), ("timestamp_secs", ctypes.c_uint64),
                ("timestamp_nanos", ctypes.c_uint32), ("signal", ctypes.c_uint64), ("elem_bytes", ctypes.c_uint32),
                ("data", ctypes.c_void_p), ("n", ctypes.c_uint64), ("has_time_lapse", ctypes.c_int32),
                ("lapse", ctypes.c_void_p), ("n_lapse", ctypes.c_uint64), ("has_prev_offset", ctypes.c_int32),
                ("prev_secs", ctypes.c_uint64), ("prev_nanos", ctypes.c_uint32), ("consumed", ctypes.c_uint64)]


P = ctypes.POINTER
vp = ctypes.c_void_p
u8p, u32p, u64p, f64p = P(ctypes.c_uint8), P(ctypes.c_uint32), P(ctypes.c_uint64), P(ctypes.c_double)
u64, u32, i32, f64, cint = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double, ctypes.c_int

# name -> argtypes (every function returns int status except buff_last_error)
SIGNATURES = {
    "buff_ctx_create": [cint, u64, P(vp)],
    "buff_ctx_destroy": [vp],
    "buff_ctx_synchronize": [vp],
    "buff_ctx_hbm_used": [vp, u64p],
    "buff_dev_alloc": [vp, u64, P(vp)],
    "buff_dev_free": [vp, vp],
    "buff_memcpy_h2d": [vp, vp, vp, u64],
    "buff_memcpy_d2h": [vp, vp, vp, u64],
    "buff_dev_fill": [vp, vp, ctypes.c_uint8, u64],
    "buff_dev_touch": [vp, vp, u64],
    "buff_dev_copy": [vp, vp, vp, u64],
    "buff_dev_read": [vp, vp, u64],
    "buff_prof_enable": [vp, cint],
    "buff_prof_reset": [vp],
    "buff_prof_query": [vp, ctypes.c_char_p, u64p, f64p],
    "buff_encode_device": [vp, vp, u64, u64, u32, P(vp)],
    "buff_encode": [vp, f64p, u64, u64, u32, P(vp)],
    "buff_chunk_info_get": [vp, P(ChunkInfo)],
    "buff_chunk_serialize": [vp, vp, u8p, u64p],
    "buff_chunk_load": [vp, u8p, u64, u64, u32, P(vp)],
    "buff_chunk_free": [vp],
    "buff_chunk_sparse_col": [vp, u32, u32p, u64p, u64p],
    "buff_filter": [vp, vp, cint, f64, u32, u32p, u64, u64p],
    "buff_filter_device": [vp, vp, cint, f64, u32, vp, vp],
    "buff_sum": [vp, vp, f64p, f64p, u64p],
    "buff_sum_ex": [vp, vp, u32, f64p, f64p, u64p],
    "buff_sum_precision": [vp, vp, i32, u32, f64p, f64p, u64p],
    "buff_sum_noalloc": [vp, vp, f64p, f64p, u64p],
    "buff_max": [vp, vp, f64p, u64p],
    "buff_max_ex": [vp, vp, u32, f64p, u64p, u64p],
    "buff_filter_cond": [vp, vp, cint, f64, u32, u32p, u32p, u64, u64p],
    "buff_filter_cond_device": [vp, vp, cint, f64, u32, vp, vp],
    "buff_materialize_cond": [vp, vp, u32, u32p, f64p, u64, u64p],
    "buff_materialize_cond_device": [vp, vp, u32, vp, vp, u64, u64p],
    "buff_max_range": [vp, vp, u64, u64, u64, u32, f64p, u64, u64p, u32p, u64],
    "buff_max_range_device": [vp, vp, u64, u64, u64, u32, vp, u64, u64p, vp],
    "buff_materialize": [vp, vp, i32, u32, f64p],
    "buff_materialize_device": [vp, vp, i32, u32, vp],
    "buff_batch_encode_device": [vp, vp, u64p, u64p, u32, u32, P(vp)],
    "buff_batch_size": [vp, u32p],
    "buff_batch_chunk": [vp, u32, P(vp)],
    "buff_batch_free": [vp],
    "buff_batch_query": [vp, vp, P(Query), u32, P(QueryResult)],
    "buff_batch_arena": [vp, P(vp), P(vp)],
    "buff_batch_tune": [vp, vp, P(Query), u32, u32p, u32p, u32p],
    "buff_batch_set_tuning": [vp, u32, u32, u32],
    "buff_tune_chunk": [vp, vp, ctypes.c_double, P(TuneChoice)],
    "buff_server_create": [vp, P(ServerCfg), P(vp)],
    "buff_server_add_chunk": [vp, vp, u32p],
    "buff_server_submit": [vp, P(Query), vp, u64p],
    "buff_server_wait": [vp, u64, P(QueryResult), f64p],
    "buff_server_poll": [vp, u64, P(cint)],
    "buff_server_release": [vp, vp],
    "buff_server_arena_stats": [vp, u64p, u64p, u64p],
    "buff_server_run_trace": [vp, P(Query), f64p, P(vp), u32, P(QueryResult), f64p],
    "buff_server_stats": [vp, u64p, u64p, u64p],
    "buff_trace_per_call": [vp, P(vp), u32, P(Query), f64p, P(vp), u32, P(QueryResult), f64p],
    "buff_server_destroy": [vp],
    "buff_csv_parse": [vp, u8p, u64, ctypes.c_uint8, u32, f64p, u64, u64p],
    "buff_csv_parse_device": [vp, vp, u64, ctypes.c_uint8, u32, vp, u64, u64p],
    "buff_construct_file_iterator_skip_newline": [ctypes.c_char_p, u32, ctypes.c_uint8, f64p, u64, u64p],
    "buff_sd_byte_fixed_encode": [P(SplitDouble), f64p, u64, u8p, u64, u64p],
    "buff_sd_buff_slice_encode": [P(SplitDouble), f64p, u64, u8p, u64, u64p],
    "buff_sd_buff_slice_decode": [P(SplitDouble), u8p, u64, f64p, u64],
    "buff_sd_buff_slice_decode_bs": [P(SplitDouble), u8p, u64, f64p, u64],
    "buff_sd_buff_slice_range_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_slice_equal_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_slice_range_filter_nosimd": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_slice_equal_filter_nosimd": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_simd_range_filter_with_slice": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_simd_equal_filter_with_slice": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_slice_buffsum": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_buff_slice_sum": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_buff_slice_buffmax": [P(SplitDouble), u8p, u64, f64p, u64p],
    "buff_sd_buff_slice_max": [P(SplitDouble), u8p, u64, f64p, u64p],
    "buff_sd_buff_simd256_encode": [P(SplitDouble), f64p, u64, u8p, u64, u64p],
    "buff_sd_byte_fixed_decode": [P(SplitDouble), u8p, u64, f64p, u64],
    "buff_sd_buff_simd256_decode": [P(SplitDouble), u8p, u64, f64p, u64],
    "buff_sd_buff_range_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_equal_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_simd_range_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_simd_equal_filter": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_byte_fixed_sum": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_byte_fixed_max": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_byte_residue_decode_with_precision": [P(SplitDouble), u8p, u64, u32, f64p, u64, u64p],
    "buff_sd_byte_residue_encode": [P(SplitDouble), f64p, u64, u8p, u64, u64p],
    "buff_sd_byte_residue_decode": [P(SplitDouble), u8p, u64, f64p, u64, u64p],
    "buff_sd_byte_residue_sum": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_byte_residue_sum_with_precision": [P(SplitDouble), u8p, u64, u32, f64p],
    "buff_sd_buff_encode_majority": [P(SplitDouble), f64p, u64, u8p, u64, u64p],
    "buff_sd_buff_decode_majority": [P(SplitDouble), u8p, u64, f64p, u64, u64p],
    "buff_sd_buff_range_filter_majority": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_equal_filter_majority": [P(SplitDouble), u8p, u64, f64, u32p, u64, u64p],
    "buff_sd_buff_sum_majority": [P(SplitDouble), u8p, u64, f64p],
    "buff_sd_buff_max_majority": [P(SplitDouble), u8p, u64, f64p],
    "buff_segment_byte_size": [P(SegmentDesc), u64p],
    "buff_segment_serialize": [P(SegmentDesc), u8p, u64, u64p],
    "buff_segment_parse": [u8p, u64, u32, P(SegmentDesc)],
    "buff_segment_encode": [vp, u8p, u64, u32, u64, u32, P(vp)],
}

_lib = None


def lib():
    """Loads libbuffgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbuffgpu.so is not built (run `python -m buff_amd.build`): " + LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        L.buff_last_error.restype = ctypes.c_char_p
        L.buff_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc):
    if rc == BUFF_OK:
        return
    msg = lib().buff_last_error().decode(errors="replace")
    cls = {BUFF_E_REF_PANIC: RefPanic, BUFF_E_UNSUPPORTED: Unsupported, BUFF_E_OOM: OutOfMemory}.get(rc, BuffError)
    raise cls(rc, msg)


def ptr(a, ct):
    return a.ctypes.data_as(P(ct))


class Chunk:
    """A resident column (buff_chunk*)."""

    def __init__(self, ctx, handle, borrowed=False):
        self.ctx = ctx
        self.handle = handle
        self.borrowed = borrowed   # owned by a Batch

    def info(self):
        inf = ChunkInfo()
        check(lib().buff_chunk_info_get(self.handle, ctypes.byref(inf)))
        return inf

    @property
    def n(self):
        return self.info().n

    def serialize(self):
        L = lib()
        n = ctypes.c_uint64(0)
        check(L.buff_chunk_serialize(self.ctx.handle, self.handle, None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint8)
        check(L.buff_chunk_serialize(self.ctx.handle, self.handle, ptr(out, ctypes.c_uint8), ctypes.byref(n)))
        return out

    def filter(self, op, pred, flags=0):
        """Returns (u32 words in BitVec layout, count)."""
        n = self.n
        words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        cnt = ctypes.c_uint64()
        opc = _opc(op)
        check(lib().buff_filter(self.ctx.handle, self.handle, opc, float(pred), flags, ptr(words, ctypes.c_uint32),
                                len(words), ctypes.byref(cnt)))
        return words[: (n + 31) // 32], cnt.value

    def filter_device(self, op, pred, words_dev, flags=0, stats_dev=None):
        opc = _opc(op)
        check(lib().buff_filter_device(self.ctx.handle, self.handle, opc, float(pred), flags, words_dev, stats_dev))

    def filter_cond(self, op, pred, cond_words, flags=0):
        """The filter over the rows set in cond_words (u32 words, the layout filter returns).
        Returns (u32 words, count)."""
        n = self.n
        cw = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        c = np.asarray(cond_words, dtype=np.uint32)[: (n + 31) // 32]
        cw[: len(c)] = c
        words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        cnt = ctypes.c_uint64()
        opc = _opc(op)
        check(lib().buff_filter_cond(self.ctx.handle, self.handle, opc, float(pred), flags, ptr(cw, ctypes.c_uint32),
                                     ptr(words, ctypes.c_uint32), len(words), ctypes.byref(cnt)))
        return words[: (n + 31) // 32], cnt.value

    def filter_cond_device(self, op, pred, cond_dev, words_dev, flags=0):
        opc = _opc(op)
        check(lib().buff_filter_cond_device(self.ctx.handle, self.handle, opc, float(pred), flags, cond_dev,
                                            words_dev))

    def max_range(self, s, e, window, flags=0):
        """buff_max_range: (window max values, argmax u32 words)."""
        n = self.n
        cap = max((e - s + window - 1) // window, 1) if window and e >= s else 1
        vals = np.zeros(cap, dtype=np.float64)
        nwin = ctypes.c_uint64()
        words = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        check(lib().buff_max_range(self.ctx.handle, self.handle, s, e, window, flags, ptr(vals, ctypes.c_double), cap,
                                   ctypes.byref(nwin), ptr(words, ctypes.c_uint32), len(words)))
        return vals[: nwin.value], words[: (n + 31) // 32]

    def max_range_device(self, s, e, window, vals_dev, cap, argmax_dev=None, flags=0):
        """Returns the number of windows (values at vals_dev)."""
        nwin = ctypes.c_uint64()
        check(lib().buff_max_range_device(self.ctx.handle, self.handle, s, e, window, flags, vals_dev, cap,
                                          ctypes.byref(nwin), argmax_dev))
        return nwin.value

    def materialize_cond(self, cond_words, flags=0):
        """Values of the rows set in cond_words (u32 words), in row order."""
        n = self.n
        cw = np.zeros(max((n + 31) // 32, 1), dtype=np.uint32)
        c = np.asarray(cond_words, dtype=np.uint32)[: (n + 31) // 32]
        cw[: len(c)] = c
        out = np.zeros(max(n, 1), dtype=np.float64)
        rows = ctypes.c_uint64()
        check(lib().buff_materialize_cond(self.ctx.handle, self.handle, flags, ptr(cw, ctypes.c_uint32),
                                          ptr(out, ctypes.c_double), n, ctypes.byref(rows)))
        return out[: rows.value]

    def materialize_cond_device(self, cond_dev, out_dev, cap, flags=0):
        """Returns the number of condition rows (their values are at out_dev)."""
        rows = ctypes.c_uint64()
        check(lib().buff_materialize_cond_device(self.ctx.handle, self.handle, flags, cond_dev, out_dev, cap,
                                                 ctypes.byref(rows)))
        return rows.value

    def sum(self, flags=0):
        """Returns (reference f64 sum, exact sum, per-sub-column u64 sums[9])."""
        r, e = ctypes.c_double(), ctypes.c_double()
        ps = np.zeros(9, dtype=np.uint64)
        check(lib().buff_sum_ex(self.ctx.handle, self.handle, flags, ctypes.byref(r), ctypes.byref(e),
                                ptr(ps, ctypes.c_uint64)))
        return r.value, e.value, ps

    def sum_noalloc(self):
        """buff_sum without device allocation (the library's static scratch): (reference sum, exact sum,
        per-sub-column u64 sums[9])."""
        r, e = ctypes.c_double(), ctypes.c_double()
        ps = np.zeros(9, dtype=np.uint64)
        check(lib().buff_sum_noalloc(self.ctx.handle, self.handle, ctypes.byref(r), ctypes.byref(e),
                                     ptr(ps, ctypes.c_uint64)))
        return r.value, e.value, ps

    def sum_precision(self, precision, flags=0):
        """byte_residue_sum_with_precision: returns (reference f64 sum, exact sum of the values truncated to
        the bits read, u64 sums read [9])."""
        r, e = ctypes.c_double(), ctypes.c_double()
        ps = np.zeros(9, dtype=np.uint64)
        check(lib().buff_sum_precision(self.ctx.handle, self.handle, int(precision), flags, ctypes.byref(r),
                                       ctypes.byref(e), ptr(ps, ctypes.c_uint64)))
        return r.value, e.value, ps

    def max(self):
        """Returns (max value as the reference prints it, max stored offset)."""
        v, o = ctypes.c_double(), ctypes.c_uint64()
        check(lib().buff_max(self.ctx.handle, self.handle, ctypes.byref(v), ctypes.byref(o)))
        return v.value, o.value

    def max_ex(self, flags=0):
        """Returns (max value, max stored offset, rows tying the max in the tail stage)."""
        v, o, a = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().buff_max_ex(self.ctx.handle, self.handle, flags, ctypes.byref(v), ctypes.byref(o), ctypes.byref(a)))
        return v.value, o.value, a.value

    def tune(self, pred):
        """Times candidate launch shapes of the range filter (at pred), sum, max and full materialization on
        this chunk and keeps the fastest for chunks of its size class. Returns one dict per operation."""
        ch = (TuneChoice * 4)()
        check(lib().buff_tune_chunk(self.ctx.handle, self.handle, float(pred), ch))
        names = ("range_filter", "sum", "max", "materialize")
        return {names[i]: {"waves_per_wg": c.waves_per_wg, "max_wg": c.max_wg, "tiles_per_wave": c.tiles_per_wave,
                           "prefetch": c.prefetch, "best_us": round(c.best_us, 2), "default_us": round(c.default_us, 2)}
                for i, c in enumerate(ch)}

    def sparse_col(self, j):
        """buff-major sparse column j: (majority byte, outlier rows, last outlier row)."""
        m, k, last = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().buff_chunk_sparse_col(self.handle, j, ctypes.byref(m), ctypes.byref(k), ctypes.byref(last)))
        return m.value, k.value, last.value

    def materialize(self, precision=-1, flags=0):
        inf = self.info()
        rows = inf.n if (precision >= 0 or flags & BUFF_SEMANTIC) else inf.decoded_rows
        out = np.zeros(max(inf.n, 1), dtype=np.float64)
        check(lib().buff_materialize(self.ctx.handle, self.handle, precision, flags, ptr(out, ctypes.c_double)))
        return out[:rows]

    def materialize_device(self, out_dev, precision=-1, flags=0):
        check(lib().buff_materialize_device(self.ctx.handle, self.handle, precision, flags, out_dev))

    def free(self):
        if self.handle and not self.borrowed:
            check(lib().buff_chunk_free(self.handle))
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


QUERY_DTYPE = np.dtype(Query)          # the same layout as buff_query
RESULT_DTYPE = np.dtype(QueryResult)   # the same layout as buff_query_result


class PreparedQueries:
    """Query descriptors already in buff_query layout (what a native caller
    holds): built once, passed to Batch.query / Server.run_trace /
    Context.trace_per_call without per-call conversion."""

    def __init__(self, qs):
        if isinstance(qs, np.ndarray) and qs.dtype == QUERY_DTYPE:
            a = np.ascontiguousarray(qs).copy()
        else:
            a = np.array([tuple(q) for q in qs], dtype=QUERY_DTYPE)
        self.n = len(a)
        self.buf = a if self.n else np.zeros(1, dtype=QUERY_DTYPE)
        self.arr = (Query * len(self.buf)).from_buffer(self.buf)   # shares the numpy memory

    def __len__(self):
        return self.n


def prepare_queries(qs):
    return qs if isinstance(qs, PreparedQueries) else PreparedQueries(qs)


def make_queries(qs):
    """qs: iterable of (op, chunk, pred, precision, flags), or PreparedQueries
    -> (ctypes array of Query, count)."""
    p = prepare_queries(qs)
    return p.arr, p.n


class QueryResults(collections.abc.Sequence):
    """The buff_query_result array of a call, read element by element (no
    per-call conversion of every result)."""

    def __init__(self, res, n):
        self._res = res
        self._n = n

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self._res[k] for k in range(*i.indices(self._n))]
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError("query result %d of %d" % (i, self._n))
        return self._res[i]


class Batch:
    """buff_batch*: chunks encoded together; queries over them run as one graph launch."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        n = ctypes.c_uint32()
        check(lib().buff_batch_size(handle, ctypes.byref(n)))
        self.chunks = []
        for i in range(n.value):
            h = ctypes.c_void_p()
            check(lib().buff_batch_chunk(handle, i, ctypes.byref(h)))
            self.chunks.append(Chunk(ctx, h.value, borrowed=True))

    def query(self, qs):
        """Returns a list of QueryResult (see include/buffgpu.h)."""
        arr, nq = make_queries(qs)
        res = (QueryResult * max(nq, 1))()
        check(lib().buff_batch_query(self.ctx.handle, self.handle, arr, nq, res))
        return QueryResults(res, nq)

    def tune(self, qs):
        arr, nq = make_queries(qs)
        g, t, f = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().buff_batch_tune(self.ctx.handle, self.handle, arr, nq, ctypes.byref(g), ctypes.byref(t),
                                    ctypes.byref(f)))
        return g.value, t.value, f.value

    def arena(self):
        b, v = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().buff_batch_arena(self.handle, ctypes.byref(b), ctypes.byref(v)))
        return b.value, v.value

    def bitmap(self, result, n):
        """The filter bitmap of one query as BitVec<u32> words (host copy)."""
        b, _ = self.arena()
        nw = (n + 63) // 64
        out = np.zeros(max(nw, 1), dtype=np.uint64)
        if nw:
            check(lib().buff_memcpy_d2h(self.ctx.handle, out.ctypes.data, b + result.out_offset, 8 * nw))
        return out.view(np.uint32)[: (n + 31) // 32]

    def values(self, result, n):
        _, v = self.arena()
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n:
            check(lib().buff_memcpy_d2h(self.ctx.handle, out.ctypes.data, v + result.out_offset, 8 * n))
        return out[:n]

    def free(self):
        if self.handle:
            for c in self.chunks:
                c.handle = None
            check(lib().buff_batch_free(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """buff_server*: continuous batching over resident chunks (see include/buffgpu.h)."""

    def __init__(self, ctx, max_batch=0, slots=0, arena_bytes=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        cfg = ServerCfg(max_batch, slots, arena_bytes)
        check(lib().buff_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value

    def add_chunk(self, chunk):
        i = ctypes.c_uint32()
        check(lib().buff_server_add_chunk(self.handle, chunk.handle, ctypes.byref(i)))
        return i.value

    def submit(self, op, chunk, pred=0.0, precision=-1, flags=0, out_dev=None):
        q = Query(op, chunk, float(pred), int(precision), int(flags))
        t = ctypes.c_uint64()
        check(lib().buff_server_submit(self.handle, ctypes.byref(q), out_dev, ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        """Returns (result, latency_us). A filter or materialization submitted without out_dev has its output
        at result.out_offset (a device address) until release(result.out_offset)."""
        r = QueryResult()
        lat = ctypes.c_double()
        check(lib().buff_server_wait(self.handle, ticket, ctypes.byref(r), ctypes.byref(lat)))
        return r, lat.value

    def poll(self, ticket):
        d = ctypes.c_int()
        check(lib().buff_server_poll(self.handle, ticket, ctypes.byref(d)))
        return bool(d.value)

    def release(self, out_addr):
        check(lib().buff_server_release(self.handle, ctypes.c_void_p(out_addr)))

    def arena_stats(self):
        """(bytes held now, most ever held, arena bytes)."""
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().buff_server_arena_stats(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def run_trace(self, qs, arrival_s, outs=None):
        """Open-loop replay: request i submitted at arrival_s[i] seconds. Returns (results, latencies_us)."""
        arr, n = make_queries(qs)
        t = np.ascontiguousarray(arrival_s, dtype=np.float64)
        res = (QueryResult * max(n, 1))()
        lat = np.zeros(max(n, 1), dtype=np.float64)
        o = None
        if outs is not None:
            o = (ctypes.c_void_p * max(n, 1))(*[ctypes.c_void_p(v) if v else None for v in outs])
        check(lib().buff_server_run_trace(self.handle, arr, ptr(t, ctypes.c_double), o, n, res,
                                          ptr(lat, ctypes.c_double)))
        return QueryResults(res, n), lat[:n]

    def stats(self):
        b, q, m = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().buff_server_stats(self.handle, ctypes.byref(b), ctypes.byref(q), ctypes.byref(m)))
        return b.value, q.value, m.value

    def destroy(self):
        if self.handle:
            check(lib().buff_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned by a context (counted against its HBM budget)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        check(lib().buff_dev_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(lib().buff_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().buff_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def fill(self, value=0):
        check(lib().buff_dev_fill(self.ctx.handle, self.ptr, value, self.nbytes))

    def touch(self):
        """Reads the buffer on the device (clean-line cache eviction)."""
        check(lib().buff_dev_touch(self.ctx.handle, self.ptr, self.nbytes))

    def copy_from(self, src, nbytes=None):
        """Device-to-device streaming copy of nbytes (default: all of src) into this buffer."""
        nbytes = src.nbytes if nbytes is None else nbytes
        if nbytes > self.nbytes or nbytes > src.nbytes:
            raise ValueError("copy of %d bytes between buffers of %d and %d" % (nbytes, src.nbytes, self.nbytes))
        check(lib().buff_dev_copy(self.ctx.handle, self.ptr, src.ptr, nbytes))

    def stream_read(self, nbytes=None):
        """Reads nbytes (default: all) with the non-temporal streaming-read kernel."""
        nbytes = self.nbytes if nbytes is None else nbytes
        if nbytes > self.nbytes:
            raise ValueError("read of %d bytes from a buffer of %d" % (nbytes, self.nbytes))
        check(lib().buff_dev_read(self.ctx.handle, self.ptr, nbytes))

    def free(self):
        if self.ptr:
            check(lib().buff_dev_free(self.ctx.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """buff_ctx*: one HIP stream on one device with an HBM budget (0 = none)."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        check(lib().buff_ctx_create(device, hbm_budget, ctypes.byref(h)))
        self.handle = h.value

    def encode(self, x, scale, flags=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().buff_encode(self.handle, ptr(x, ctypes.c_double), len(x), scale, flags, ctypes.byref(h)))
        return Chunk(self, h.value)

    def encode_device(self, x_dev, n, scale, flags=0):
        h = ctypes.c_void_p()
        check(lib().buff_encode_device(self.handle, x_dev, n, scale, flags, ctypes.byref(h)))
        return Chunk(self, h.value)

    def encode_segment(self, seg_bytes, scale, flags=0, elem_bytes=8):
        """byte_fixed_encode on a serialized Segment<T> (buff_segment_encode)."""
        b = np.frombuffer(bytes(seg_bytes), dtype=np.uint8)
        h = ctypes.c_void_p()
        check(lib().buff_segment_encode(self.handle, ptr(b, ctypes.c_uint8) if len(b) else None, len(b), elem_bytes,
                                        scale, flags, ctypes.byref(h)))
        return Chunk(self, h.value)

    def csv_parse(self, text, delim=b",", skip_val=0):
        """construct_file_iterator_skip_newline::<f64> over `text` (bytes): the values."""
        t = np.frombuffer(bytes(text), dtype=np.uint8)
        cap = len(t) // 2 + 1
        out = np.zeros(max(cap, 1), dtype=np.float64)
        n = ctypes.c_uint64()
        check(lib().buff_csv_parse(self.handle, ptr(t, ctypes.c_uint8) if len(t) else None, len(t), ord(delim),
                                   skip_val, ptr(out, ctypes.c_double), cap, ctypes.byref(n)))
        return out[: n.value]

    def csv_parse_device(self, text_dev, length, out_dev, cap, delim=b",", skip_val=0):
        """Parses text already in device memory into device memory; returns the value count."""
        n = ctypes.c_uint64()
        check(lib().buff_csv_parse_device(self.handle, text_dev, length, ord(delim), skip_val, out_dev, cap,
                                          ctypes.byref(n)))
        return n.value

    def load(self, stream, scale, flags=0):
        s = np.ascontiguousarray(stream, dtype=np.uint8)
        h = ctypes.c_void_p()
        check(lib().buff_chunk_load(self.handle, ptr(s, ctypes.c_uint8), len(s), scale, flags, ctypes.byref(h)))
        return Chunk(self, h.value)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def trace_per_call(self, chunks, qs, arrival_s, outs=None):
        """Open-loop replay of requests as plain single-chunk calls, one at a time (no batching).
        Returns (results, latencies_us)."""
        arr, n = make_queries(qs)
        hs = (ctypes.c_void_p * max(len(chunks), 1))(*[c.handle for c in chunks])
        t = np.ascontiguousarray(arrival_s, dtype=np.float64)
        res = (QueryResult * max(n, 1))()
        lat = np.zeros(max(n, 1), dtype=np.float64)
        o = None
        if outs is not None:
            o = (ctypes.c_void_p * max(n, 1))(*[ctypes.c_void_p(v) if v else None for v in outs])
        check(lib().buff_trace_per_call(self.handle, hs, len(chunks), arr, ptr(t, ctypes.c_double), o, n, res,
                                        ptr(lat, ctypes.c_double)))
        return QueryResults(res, n), lat[:n]

    def set_batch_tuning(self, grid, tiles_per_item, first_planes):
        """Pins the batch shape (grid, tiles per item, 1 or 8 first filter planes)
        for later batches and for servers created afterwards."""
        check(lib().buff_batch_set_tuning(self.handle, grid, tiles_per_item, first_planes))

    def batch_encode_device(self, x_dev, row_off, scales, flags=0):
        ro = np.ascontiguousarray(row_off, dtype=np.uint64)
        sc = np.ascontiguousarray(scales, dtype=np.uint64)
        h = ctypes.c_void_p()
        check(lib().buff_batch_encode_device(self.handle, x_dev, ptr(ro, ctypes.c_uint64), ptr(sc, ctypes.c_uint64),
                                             len(sc), flags, ctypes.byref(h)))
        return Batch(self, h.value)

    def synchronize(self):
        check(lib().buff_ctx_synchronize(self.handle))

    def hbm_used(self):
        v = ctypes.c_uint64()
        check(lib().buff_ctx_hbm_used(self.handle, ctypes.byref(v)))
        return v.value

    def prof_enable(self, on=True):
        check(lib().buff_prof_enable(self.handle, int(on)))

    def prof_reset(self):
        check(lib().buff_prof_reset(self.handle))

    def prof_query(self, kernel):
        n, ms = ctypes.c_uint64(), ctypes.c_double()
        check(lib().buff_prof_query(self.handle, kernel.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def destroy(self):
        if self.handle:
            check(lib().buff_ctx_destroy(self.handle))
            self.handle = None


def _segment_desc(seg, keep):
    """A SegmentDesc over the fields of `seg` (a buff_amd.segment.Segment); `keep`
    collects the buffers the descriptor points into."""
    d = SegmentDesc()
    d.method = seg.method
    d.dictionary_id = seg.dictionary_id
    model = bytes(seg.model)
    mb = ctypes.create_string_buffer(model, max(len(model), 1))
    keep.append(mb)
    d.model = ctypes.cast(mb, ctypes.c_void_p) if model else None
    d.model_len = len(model)
    d.timestamp_secs, d.timestamp_nanos = seg.timestamp
    d.signal = seg.signal
    d.elem_bytes = seg.elem_bytes
    arr = np.ascontiguousarray(seg.data, dtype="<f8" if seg.elem_bytes == 8 else "<f4")
    keep.append(arr)
    d.data = arr.ctypes.data if len(arr) else None
    d.n = len(arr)
    if seg.time_lapse is not None:
        rec = np.zeros(len(seg.time_lapse), dtype=[("secs", "<u8"), ("nanos", "<u4")])
        for i, (sec, ns) in enumerate(seg.time_lapse):
            rec[i] = (sec, ns)
        keep.append(rec)
        d.has_time_lapse = 1
        d.lapse = rec.ctypes.data if len(rec) else None
        d.n_lapse = len(rec)
    if seg.prev_seg_offset is not None:
        d.has_prev_offset = 1
        d.prev_secs, d.prev_nanos = seg.prev_seg_offset
    return d


def segment_serialize(seg):
    """Segment::convert_to_bytes (buff_segment_serialize)."""
    keep = []
    d = _segment_desc(seg, keep)
    n = ctypes.c_uint64()
    check(lib().buff_segment_serialize(ctypes.byref(d), None, 0, ctypes.byref(n)))
    out = np.zeros(max(n.value, 1), dtype=np.uint8)
    check(lib().buff_segment_serialize(ctypes.byref(d), ptr(out, ctypes.c_uint8), n.value, ctypes.byref(n)))
    return bytes(out[: n.value])


def segment_byte_size(seg):
    """Segment::get_byte_size (buff_segment_byte_size)."""
    keep = []
    d = _segment_desc(seg, keep)
    n = ctypes.c_uint64()
    check(lib().buff_segment_byte_size(ctypes.byref(d), ctypes.byref(n)))
    return n.value


def segment_parse(b, elem_bytes=8):
    """Segment::convert_from_bytes (buff_segment_parse): the fields as a dict."""
    buf = np.frombuffer(bytes(b), dtype=np.uint8)
    d = SegmentDesc()
    check(lib().buff_segment_parse(ptr(buf, ctypes.c_uint8) if len(buf) else None, len(buf), elem_bytes,
                                   ctypes.byref(d)))
    base = buf.ctypes.data

    def at(p, k):
        return bytes(buf[p - base: p - base + k]) if k else b""

    dt = np.dtype("<f8" if elem_bytes == 8 else "<f4")
    data = np.frombuffer(at(d.data, d.n * elem_bytes), dtype=dt).copy() if d.n else np.zeros(0, dt)
    lapse = None
    if d.has_time_lapse:
        raw = at(d.lapse, 12 * d.n_lapse)
        lapse = [(int.from_bytes(raw[12 * i: 12 * i + 8], "little"), int.from_bytes(raw[12 * i + 8: 12 * i + 12],
                                                                                    "little"))
                 for i in range(d.n_lapse)]
    return {"method": d.method, "dictionary_id": d.dictionary_id,
            "model": at(d.model, d.model_len) if d.model_len else b"",
            "ts_secs": d.timestamp_secs, "ts_nanos": d.timestamp_nanos, "signal": d.signal, "data": data,
            "time_lapse": lapse, "prev_offset": (d.prev_secs, d.prev_nanos) if d.has_prev_offset else None,
            "consumed": d.consumed}
