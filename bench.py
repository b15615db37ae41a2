"""BUFF hot path on one MI355X: encode, range filters, sum, max, materialize.

Workload (one step): the 268,435,456-value job of two float64 columns of
134,217,728 rows each, resident in HBM before timing starts:
  walk    — random walk x_{t+1} = x_t + N(0,1), 4 decimals, scale 10^4
  uniform — U[-1000, 1000], 2 decimals, scale 10^2
A step encodes both columns, runs range filters (x > pred) at 0.1%, 10% and
50% selectivity on each, sums and maxes both, and materializes the walk at 4
and at 2 decimals and the uniform column at 2 decimals (its 4-decimal request
is a reference panic). `value` is GB/s of float64 input: 8 bytes per row per
operation over the step time.

Per-kernel durations come from HIP events attached to the kernel launches
(start of the operation's first kernel to end of its last) during a repetition
of the timed steps. `roofline` is reported for the kernel with the largest
share of the step, from its algorithmic bytes (HBM bytes an ideal kernel must
move) over its average duration against 8 TB/s. A cold-cache pass (the
memory-side cache refilled with 1 GiB of other, clean data before each launch)
is reported beside it. `filter_project` times the conditional filter and the
gather of a filter's rows (buff_range_filter_condition, buff_decode_condition)
on the walk column, cold; `window_max` times buff_max_range over the walk
column at two window sizes; `slice` times the walk column as a byte-slice
chunk (buff_slice_encode layout): encode, filters, decode, sum, stream write;
`ingest` parses the reference's CSV data file (repeated 40 times, resident in
HBM) into f64 values on the device.

Further sections of the same JSON line (each skippable):
  config3 — an outlier-heavy column of 67,108,864 rows in buff and buff-major:
            per-operation device time and roofline fractions;
  config4 — 4096 chunks of uneven length and precision: one-launch batch
            encode, 16,384 mixed queries as one graph launch against one call
            per query;
  config5 — chunks encoded until 48 GB are resident, a request mix served by
            the native server with continuous batching against one launch per
            request, saturated and under Poisson arrivals (p50/p99).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N] [--no-cpu]
       [--no-config3] [--no-config4] [--no-config5]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "range-filter/sum/encode GB/s vs HBM roofline; batched-query p50/p99 latency"


TRAFFIC_FILE = "profiles/pmc_traffic_bench_268M.json"
TRAFFIC_ROWS = 1 << 27
# kernel family of the bench's profiler names -> prefix of the rocprofv3 names
TRAFFIC_FAMILY = {"k_materialize": "void k_materialize<", "k_sum": "void k_sum<", "k_filter_gt": "void k_filter<0",
                  "k_encode": "k_encode", "k_minmax": "k_minmax", "k_max": "void k_max_top<"}


def pmc_traffic(kernel, n):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated), measured by rocprofv3 --pmc on
    this same workload and committed beside the kernel-trace summary; None for other row counts."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if n != TRAFFIC_ROWS or kernel not in TRAFFIC_FAMILY or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)["kernels"]
    pre = TRAFFIC_FAMILY[kernel]
    shapes = [v for k, v in d.items() if k == pre or (pre.endswith("<") or pre.endswith("<0")) and k.startswith(pre)]
    if not shapes:
        return None
    return round(1024 * sum(v["read_kib"] + v["write_kib"] for v in shapes) / len(shapes))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1 << 27, help="rows per column (default 134,217,728)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--cpu-rows", type=int, default=1 << 25)
    ap.add_argument("--no-config3", action="store_true", help="skip the outlier-heavy (buff vs buff-major) section")
    ap.add_argument("--config3-rows", type=int, default=1 << 26)
    ap.add_argument("--no-residue", action="store_true", help="skip the RAPG residue-format section")
    ap.add_argument("--residue-rows", type=int, default=1 << 26)
    ap.add_argument("--no-config4", action="store_true", help="skip the batched-form section")
    ap.add_argument("--config4-chunks", type=int, default=4096)
    ap.add_argument("--config4-queries", type=int, default=16384)
    ap.add_argument("--no-config5", action="store_true", help="skip the served-form section")
    ap.add_argument("--config5-gb", type=float, default=48.0, help="resident compressed bytes (GB, 1e9)")
    ap.add_argument("--config5-requests", type=int, default=3000, help="requests per load point")
    ap.add_argument("--config5-slots", type=int, default=2, help="batches in flight (continuous batching)")
    ap.add_argument("--config5-max-batch", type=int, default=4096, help="requests per batch (continuous batching)")
    return ap.parse_args()


class Column:
    def __init__(self, name, x, scale, mat_precisions):
        self.name = name
        self.n = len(x)
        self.scale = scale
        self.mat_precisions = mat_precisions
        self.preds = {sel: float(np.quantile(x, 1.0 - sel, method="lower")) for sel in (0.001, 0.10, 0.50)}


def algorithmic_bytes(info, op, stats=None, precision=-1):
    """HBM bytes an ideal implementation of one launch must move."""
    n, F, P, r = info.n, info.width, info.planes, info.tail_bits
    if op == "k_minmax":
        return 8.0 * n
    if op == "k_encode":
        return 8.0 * n + n * F / 8.0
    if op == "k_filter_gt":
        b = n * (1 + 1 / 8.0)
        if stats is not None:
            b += float(sum(int(stats[k]) for k in range(1, P))) + float(stats[P]) * r / 8.0
        return b
    if op == "k_sum":
        return n * (P + r / 8.0)
    if op == "k_max":
        return float(n)
    if op == "k_materialize":
        bits = F if precision < 0 else precision
        return n * (bits / 8.0 + 8.0)
    raise KeyError(op)


PRECISION_MAP = {1: 5, 2: 8, 3: 11, 4: 15, 5: 18, 6: 21}


def precision_R(info, p):
    """Stored bits a reduced-precision read keeps (split_double.rs:1642-1664)."""
    if p < 0:
        return info.width
    ilen, dlen = info.ilen, info.dlen
    bits = PRECISION_MAP[p]
    if bits > dlen:
        return info.width
    R = (ilen + bits) & 0xFFFFFFFF
    if R // 8 < ((dlen + ilen) & 0xFFFFFFFF) // 8 and R % 8:
        R = (R // 8 + 1) * 8
    return min(R, info.width)


def _fixed(x, dlen):
    return np.trunc(np.ldexp(x, dlen)).astype(np.int64)


def stored_offsets(x, info):
    """The offsets a plain chunk stores: (trunc(x * 2^dlen) - base) mod 2^F (prec_double.rs:190-205;
    split_double.rs:369-414 keeps the low F bits)."""
    F = info.width
    off = (_fixed(x, info.dlen) - info.base).astype(np.uint64)
    return off & np.uint64((1 << F) - 1) if F < 64 else off


def u64_sum(a):
    """Exact sum of a uint64 array (summed in two 32-bit halves, so no partial sum wraps)."""
    a = np.asarray(a, dtype=np.uint64)
    return (int(np.sum(a >> np.uint64(32), dtype=np.uint64)) << 32) + int(np.sum(a & np.uint64(0xFFFFFFFF),
                                                                                 dtype=np.uint64))


def _words64(mask):
    b = np.packbits(mask, bitorder="little")
    return np.concatenate([b, np.zeros((-len(b)) % 8, np.uint8)]).view(np.uint64)


def filter_line_bytes(off, info, fp):
    """HBM bytes a range filter must move at 128-byte line granularity: every
    line of the top plane and of the result bitmap, then on each later plane the
    lines holding a row still tied with the target (128 consecutive rows share a
    line of a plane), and the tail bit-planes of the tiles with tied rows left
    (one line per bit-plane and 1024-row tile)."""
    n, F, P, r = info.n, info.width, info.planes, info.tail_bits
    ntiles = (n + 1023) // 1024
    b = ntiles * 1024.0 * (1 if P else 0) + ((n + 63) // 64) * 8.0
    t = fp - info.base
    if t < 0 or not P:
        return b
    t &= (1 << F) - 1
    cand = None
    for p in range(P):
        sh = np.uint64(F - 8 * (p + 1))
        tb = np.uint64((t >> int(sh)) & 0xFF)
        if cand is None:
            cand = np.flatnonzero(((off >> sh) & np.uint64(0xFF)) == tb)
        else:
            b += 128.0 * len(np.unique(cand // 128))
            cand = cand[((off[cand] >> sh) & np.uint64(0xFF)) == tb]
    if r and len(cand):
        b += 128.0 * r * len(np.unique(cand // 1024))
    return b


def verify_plain(ctx, ch, x, preds, mat_precisions, words, out, buff_amd):
    """Checks a plain chunk's answers against numpy on its stored offsets: the
    base, range-filter bitmaps and counts at each predicate, the exact sum and
    the reference fold (within 1e-12 of the summed magnitudes; without the
    base term when F % 8 == 0), the max, and the materialized doubles at each
    precision (BUFF_SEMANTIC: the top R stored bits). Returns the mismatches."""
    from fractions import Fraction

    bad = []
    info = ch.info()
    n, F, dlen, mn = info.n, info.width, info.dlen, info.base
    f = _fixed(x, dlen)
    if mn != int(f.min()):
        bad.append("base")
    off = stored_offsets(x, info)
    del f
    nw = (n + 63) // 64
    for pred in preds:
        fp = int(_fixed(np.array([pred]), dlen)[0])
        if fp < mn:
            m = np.ones(n, bool)
        else:
            t = fp - mn if F < 8 else (fp - mn) & ((1 << F) - 1)
            m = off > np.uint64(t)
        ch.filter_device("range", pred, words.ptr)
        if not np.array_equal(words.download(np.uint64, nw), _words64(m)[:nw]):
            bad.append("range filter at %r" % pred)
    so = u64_sum(off)
    ref, exact, _ = ch.sum()
    if exact != float(Fraction(n * mn + so, 1 << dlen)):
        bad.append("exact sum")
    want = float(Fraction((0 if (F >= 8 and F % 8 == 0) else n * mn) + so, 1 << dlen))
    if abs(ref - want) > 1e-12 * (n * abs(mn) + so) / 2.0 ** dlen:
        bad.append("reference sum")
    v, o = ch.max()
    mo = int(off.max()) if n else 0
    if o != mo or v != float(mo + mn) / 2.0 ** dlen:
        bad.append("max")
    for p in mat_precisions:
        ch.materialize_device(out.ptr, p, flags=buff_amd.BUFF_SEMANTIC)
        R = precision_R(info, p)
        sh = np.uint64(F - R)
        want_v = (((off >> sh) << sh).astype(np.int64) + mn).astype(np.float64) / 2.0 ** dlen
        if not np.array_equal(out.download(np.float64, n).view(np.uint64), want_v.view(np.uint64)):
            bad.append("materialize at precision %d" % p)
    return bad


def verify_major(ch, x, preds, words, out, buff_amd):
    """Checks a buff-major chunk's intended answers (BUFF_SEMANTIC) against numpy
    on trunc(x * 2^dlen): filter bitmaps and counts, the exact sum, the max and
    the full decode. Returns the mismatches."""
    from fractions import Fraction

    bad = []
    info = ch.info()
    n, dlen = info.n, info.dlen
    f = _fixed(x, dlen)
    nw = (n + 63) // 64
    for op, pred in preds:
        fp = int(_fixed(np.array([pred]), dlen)[0])
        m = f > fp if op == "range" else f == fp
        ch.filter_device(op, pred, words.ptr, flags=buff_amd.BUFF_SEMANTIC)
        if not np.array_equal(words.download(np.uint64, nw), _words64(m)[:nw]):
            bad.append("%s filter at %r" % (op, pred))
    _, exact, _ = ch.sum(flags=buff_amd.BUFF_SEMANTIC)
    fmin = int(f.min())
    if exact != float(Fraction(n * fmin + u64_sum((f - fmin).astype(np.uint64)), 1 << dlen)):
        bad.append("exact sum")
    if ch.max_ex(flags=buff_amd.BUFF_SEMANTIC)[0] != float(f.max()) / 2.0 ** dlen:
        bad.append("max")
    ch.materialize_device(out.ptr, -1, flags=buff_amd.BUFF_SEMANTIC)
    if not np.array_equal(out.download(np.float64, n).view(np.uint64),
                          (f.astype(np.float64) / 2.0 ** dlen).view(np.uint64)):
        bad.append("decode")
    return bad


def major_range_count(off, info, majors, fp):
    """Rows buff_range_filter_majority (buff_simd.rs:863-1277) reports for a
    fixed-point predicate fp, from the stored offsets of a buff-major chunk:
    sparse column 0 as a plain byte; a later sparse column adds its outliers
    above the target byte, and its non-outlier rows when the majority byte is
    above it (target 0 with majority 0 keeps the candidates' outliers and ends
    the scan); then the dense planes and the tail on the remaining candidates."""
    n, F, S, base = info.n, info.width, info.sparse_cols, info.base
    if fp < base:
        return n
    t = fp - base
    cand = np.ones(n, bool)
    res = np.zeros(n, bool)
    for j in range(S):
        sh = np.uint64(F - 8 * (j + 1))
        b = (off >> sh) & np.uint64(0xFF)
        tb = np.uint64((t >> int(sh)) & 0xFF)
        mask = b != np.uint64(majors[j])
        if j == 0:
            res |= cand & (b > tb)
            cand &= b == tb
        elif tb == 0 and majors[j] == 0:
            res |= cand & mask
            cand[:] = False
        else:
            res |= mask & (b > tb)
            if majors[j] > tb:
                res |= cand & ~mask
            cand &= b == tb
    for p in range(info.planes):
        sh = np.uint64(F - 8 * (S + p + 1))
        b = (off >> sh) & np.uint64(0xFF)
        tb = np.uint64((t >> int(sh)) & 0xFF)
        res |= cand & (b > tb)
        cand &= b == tb
    r = info.tail_bits
    if r:
        lowm = np.uint64((1 << r) - 1)
        res |= cand & ((off & lowm) > np.uint64(t & ((1 << r) - 1)))
    return int(np.count_nonzero(res))


def major_filter_rows(x, info, majors, pred, dlen):
    """Rows a buff-major range filter must touch: its candidates entering each
    dense plane and the tail (the compat rules of buff_simd.rs:863-1277)."""
    base, F, S = info.base, info.width, info.sparse_cols
    off = (_fixed(x, dlen) - base).astype(np.uint64)
    t = int(_fixed(np.array([pred]), dlen)[0]) - base
    if t < 0:
        return 0, 0
    cand = np.ones(len(x), bool)
    for j in range(S):
        sh = np.uint64(F - 8 * (j + 1))
        b = ((off >> sh) & np.uint64(0xFF)).astype(np.int64)
        tb = (t >> int(sh)) & 0xFF
        if j > 0 and tb == 0 and majors[j] == 0:
            cand[:] = False
        else:
            cand &= b == tb
    rows = 0
    for p in range(info.planes):
        rows += int(cand.sum())
        sh = np.uint64(F - 8 * (S + p + 1))
        cand &= ((off >> sh) & np.uint64(0xFF)).astype(np.int64) == ((t >> int(sh)) & 0xFF)
    return rows, int(cand.sum()) if info.tail_bits else 0


def bench_filter_project(ctx, col, ch, info, words, out, flush, buff_amd):
    """Filter -> project on the walk column, each launch after a 1 GiB cache
    eviction (HIP events, mean of 3): the range filter at 10% restricted to an
    earlier filter's rows (x above the median: the condition holds half the
    rows), and the values of the rows of a 10% and a 0.1% filter gathered in row
    order (BUFF_SEMANTIC). Algorithmic bytes: condition bitmap + result bitmap +
    top plane for the filter; condition bitmap + every plane and tail bit of the
    tiles holding condition rows + 8 bytes per value for the gather."""
    n = info.n
    nw = (n + 63) // 64
    cond = ctx.alloc(nw * 8)
    res = {}

    def timed(fn, kname, trials=3):
        ts = []
        for _ in range(trials):
            ctx.prof_reset()
            flush.touch()
            ctx.synchronize()
            ctx.prof_enable(True)
            fn()
            ctx.synchronize()
            ctx.prof_enable(False)
            launches, ms = ctx.prof_query(kname)
            ts.append(ms * 1e3 / max(launches, 1))
        return float(np.mean(ts))

    def entry(us, b, extra):
        g = b / (us * 1e-6) / 1e9
        d = {"avg_us": round(us, 2), "algorithmic_bytes": round(b), "achieved_GBps": round(g, 1),
             "frac": round(g / HBM_PEAK_GBPS, 4)}
        d.update(extra)
        return d

    ch.filter_device("range", col.preds[0.50], cond.ptr)
    us = timed(lambda: ch.filter_cond_device("range", col.preds[0.10], cond.ptr, words.ptr), "k_filter_cond_gt")
    res["filter_10pct_within_50pct"] = entry(us, n / 8.0 + n / 8.0 + n, {})
    per_row = info.planes + info.tail_bits / 8.0
    for sel in (0.10, 0.001):
        ch.filter_device("range", col.preds[sel], cond.ptr)
        bits = cond.download(np.uint64, nw)
        rows = int(np.unpackbits(bits.view(np.uint8)).sum())
        tb = np.concatenate([bits, np.zeros((-nw) % 16, np.uint64)]).reshape(-1, 16)   # 16 words = one tile
        tiles = int(np.count_nonzero(tb.any(axis=1)))
        us = timed(lambda: ch.materialize_cond_device(cond.ptr, out.ptr, n, flags=buff_amd.BUFF_SEMANTIC),
                   "k_cond_decode")
        b = n / 8.0 + tiles * 1024 * per_row + 8.0 * rows
        res["gather_%g%%" % (sel * 100)] = entry(us, b, {"rows": rows, "tiles_with_rows": tiles,
                                                         "GBps_f64_out": round(8.0 * rows / (us * 1e-6) / 1e9, 1)})
    cond.free()
    return res


def bench_window_max(ctx, col, ch, info, words, out, flush, buff_amd):
    """buff_max_range on the whole walk column, each launch after a 1 GiB cache
    eviction (HIP events, mean of 3), with the argmax bitmap (BUFF_SEMANTIC):
    windows of 1000 and of 2^20 rows. Algorithmic bytes: every plane and tail
    bit once, the argmax bitmap and 8 bytes per window maximum."""
    n = info.n
    res = {}
    for w in (1000, 1 << 20):
        ts = []
        k = 0
        for _ in range(3):
            ctx.prof_reset()
            flush.touch()
            ctx.synchronize()
            ctx.prof_enable(True)
            k = ch.max_range_device(0, n, w, out.ptr, n, argmax_dev=words.ptr, flags=buff_amd.BUFF_SEMANTIC)
            ctx.synchronize()
            ctx.prof_enable(False)
            launches, ms = ctx.prof_query("k_wmax")
            ts.append(ms * 1e3 / max(launches, 1))
        us = float(np.mean(ts))
        b = n * (info.planes + info.tail_bits / 8.0) + n / 8.0 + 8.0 * k
        g = b / (us * 1e-6) / 1e9
        res["window_%d" % w] = {"windows": k, "avg_us": round(us, 2), "algorithmic_bytes": round(b),
                                "achieved_GBps": round(g, 1), "frac": round(g / HBM_PEAK_GBPS, 4)}
    return res


def bench_slice(ctx, col, xdev, words, out, flush, buff_amd):
    """The walk column as a byte-slice chunk (BUFF_SLICE), each launch after a
    1 GiB cache eviction (HIP events, mean of 3): encode, the AVX2-form range
    filter at 10% and equality on a stored value, full decode, sum, and the
    slice stream kernels (planes XOR 0x80 plus the right-padded last plane).
    The resident form is byte_fixed's, so these are its kernels with the slice
    targets; algorithmic bytes as for byte_fixed, the stream write n*ceil(F/8)."""
    n = col.n
    ch = ctx.encode_device(xdev.ptr, n, col.scale, flags=buff_amd.BUFF_SLICE)
    info = ch.info()
    per_row = info.planes + info.tail_bits / 8.0
    ns = max(1, -(-info.width // 8))

    def timed(fn, kname, trials=3):
        ts = []
        for _ in range(trials):
            ctx.prof_reset()
            flush.touch()
            ctx.synchronize()
            ctx.prof_enable(True)
            r = fn()
            ctx.synchronize()
            ctx.prof_enable(False)
            if r is not None and hasattr(r, "free"):
                r.free()
            launches, ms = ctx.prof_query(kname)
            ts.append(ms * 1e3 / max(launches, 1))
        return float(np.mean(ts))

    def entry(us, b):
        g = b / (us * 1e-6) / 1e9
        return {"avg_us": round(us, 2), "algorithmic_bytes": round(b), "achieved_GBps": round(g, 1),
                "frac": round(g / HBM_PEAK_GBPS, 4), "GBps_f64_input": round(8.0 * n / (us * 1e-6) / 1e9, 1)}

    res = {"rows": n, "F": info.width, "stream_bytes": info.stream_bytes}
    res["encode"] = entry(timed(lambda: ctx.encode_device(xdev.ptr, n, col.scale, flags=buff_amd.BUFF_SLICE),
                                "k_encode"), 8.0 * n + n * info.width / 8.0)
    res["range_10pct"] = entry(timed(lambda: ch.filter_device("range", col.preds[0.10], words.ptr), "k_filter_gt"),
                               n * (1 + 1 / 8.0))
    res["equal_stored_value"] = entry(timed(lambda: ch.filter_device("equal", col.preds[0.50], words.ptr),
                                            "k_filter_eq"), n * (1 + 1 / 8.0))
    res["decode"] = entry(timed(lambda: ch.materialize_device(out.ptr, -1), "k_materialize"),
                          n * (info.width / 8.0 + 8.0))
    res["sum"] = entry(timed(lambda: ch.sum(), "k_sum"), n * per_row)
    res["stream_write"] = entry(timed(lambda: ch.serialize(), "k_serialize"), n * per_row + n * ns)
    ch.free()
    return res


def bench_ingest(ctx, flush):
    """CSV ingest (construct_file_iterator_skip_newline::<f64>) on the device: the
    reference's data file (500 lines x 1000 items) repeated 40 times, resident
    in HBM, parsed into f64 values, each run after a 1 GiB cache eviction (HIP
    events, mean of 3). Algorithmic bytes: the text once plus 8 bytes per
    value; the kernels read the text twice (the item scan and the parse)."""
    import gzip
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                        "randomwalkdatasample1k-1k.csv.gz")
    one = gzip.open(path).read()
    text = b"\n".join([one] * 40)
    n_text = len(text)
    tb = ctx.alloc(n_text)
    tb.upload(np.frombuffer(text, dtype=np.uint8))
    cap = 40 * 500 * 1000
    vb = ctx.alloc(cap * 8)
    kernels = ("k_csv_count", "k_csv_parse")
    ts = {k: [] for k in kernels}
    n = 0
    for _ in range(3):
        ctx.prof_reset()
        flush.touch()
        ctx.synchronize()
        ctx.prof_enable(True)
        n = ctx.csv_parse_device(tb.ptr, n_text, vb.ptr, cap)
        ctx.synchronize()
        ctx.prof_enable(False)
        for k in kernels:
            launches, ms = ctx.prof_query(k)
            ts[k].append(ms * 1e3 / max(launches, 1))
    us = {k: float(np.mean(v)) for k, v in ts.items()}
    total = sum(us.values())
    b = n_text + 8.0 * n
    g = b / (total * 1e-6) / 1e9
    res = {"text_bytes": n_text, "values": n, "kernel_us": {k: round(v, 2) for k, v in us.items()},
           "avg_us": round(total, 2), "algorithmic_bytes": round(b), "achieved_GBps": round(g, 1),
           "frac": round(g / HBM_PEAK_GBPS, 4), "text_GBps": round(n_text / (total * 1e-6) / 1e9, 1),
           "values_per_s": round(n / (total * 1e-6), 1)}
    tb.free()
    vb.free()
    return res


def bench_config3(ctx, args, buff_amd, datagen):
    """Outlier-heavy column (95% N(100, 0.5), 5% +-10^U[6,9], 3 decimals, seed 4) in
    the plain and the buff-major layout: encode, range filter at 10%, equality on a
    majority value and on an outlier value, sum, max, full decode. Per-op device
    time from the library's HIP events; GB/s of float64 input = 8 n / time."""
    n = args.config3_rows
    scale, dlen = 10 ** 3, 11
    x = datagen.outlier_heavy(n, 3, 4)
    xb = ctx.alloc(x.nbytes)
    xb.upload(x)
    words = ctx.alloc(((n + 63) // 64) * 8)
    out = ctx.alloc(n * 8)
    pred10 = datagen.percentile_pred(x, 0.10)
    in_major = float(x[np.argsort(np.abs(x - 100.0))[n // 3]])
    outlier = float(x[int(np.flatnonzero(np.abs(x) > 1e5)[7])])
    reps = 5
    res = {"workload": "outlier-heavy column, %d float64 rows (95%% N(100,0.5), 5%% +-10^U[6,9]), 3 decimals" % n,
           "rows": n, "predicates": {"range_10pct": pred10, "equal_majority": in_major, "equal_outlier": outlier}}
    for fmt, flags in (("buff", 0), ("buff-major", buff_amd.BUFF_MAJORITY)):
        ops = {
            "encode": lambda: ctx.encode_device(xb.ptr, n, scale, flags=flags),
        }
        ch = ops["encode"]()
        info = ch.info()
        majors = [ch.sparse_col(j)[0] for j in range(info.sparse_cols)]
        n_out = [ch.sparse_col(j)[1] for j in range(info.sparse_cols)]
        queries = {
            "range_10pct": lambda: ch.filter_device("range", pred10, words.ptr),
            "equal_majority": lambda: ch.filter_device("equal", in_major, words.ptr),
            "equal_outlier": lambda: ch.filter_device("equal", outlier, words.ptr),
            "sum": lambda: ch.sum(),
            "max": lambda: ch.max(),
            "decode": lambda: ch.materialize_device(out.ptr, -1),
        }
        fres = {"F": info.width, "sparse_cols": info.sparse_cols, "majors": majors, "outliers": n_out,
                "planes": info.planes, "tail_bits": info.tail_bits, "device_bytes": info.device_bytes, "ops": {}}
        for name, fn in [("encode", None)] + list(queries.items()):
            ctx.synchronize()
            ctx.prof_reset()
            ctx.prof_enable(True)
            t = time.perf_counter()
            for _ in range(reps):
                if name == "encode":
                    ops["encode"]().free()
                else:
                    fn()
            ctx.synchronize()
            wall = (time.perf_counter() - t) / reps
            ctx.prof_enable(False)
            kern = {}
            for k in ("k_minmax", "k_encode", "k_major_encode", "k_major_scan", "k_major_place", "k_filter_gt",
                      "k_filter_eq", "k_major_filter_gt", "k_major_filter_eq", "k_fill_bitmap", "k_sum",
                      "k_major_sum", "k_max", "k_major_max", "k_materialize", "k_major_decode"):
                l, ms = ctx.prof_query(k)
                if l:
                    kern[k] = round(ms / reps * 1e3, 2)
            dev_us = sum(kern.values())
            fres["ops"][name] = {"device_us": round(dev_us, 2), "wall_us": round(wall * 1e6, 1),
                                 "GBps_f64_input": round(8.0 * n / (dev_us * 1e-6) / 1e9, 1) if dev_us else None,
                                 "kernels_us": kern}
        # algorithmic bytes of the two ops the headline metric names
        S, P, r = info.sparse_cols, info.planes, info.tail_bits
        # sparse columns as the filter reads them: per outlier its byte and its
        # u16 row in the tile, per tile the prefix count
        sparse_b = 3.0 * sum(n_out) + S * ((n + 1023) // 1024) * 4.0
        if fmt == "buff-major":
            rows_p, rows_t = major_filter_rows(x, info, majors, pred10, dlen)
            fb = sparse_b + rows_p + rows_t * r / 8.0 + n / 8.0
            sb = sum(n_out) + n * (P + r / 8.0)
        else:
            stats = ctx.alloc(9 * 8)
            stats.fill(0)
            ch.filter_device("range", pred10, words.ptr, stats_dev=stats.ptr)
            st = stats.download(np.uint64, 9)
            fb = n * (1 + 1 / 8.0) + float(sum(int(st[k]) for k in range(1, P))) + float(st[P]) * r / 8.0
            sb = n * (P + r / 8.0)
            stats.free()
        for name, b in (("range_10pct", fb), ("sum", sb)):
            us = fres["ops"][name]["device_us"]
            if us:
                g = b / (us * 1e-6) / 1e9
                fres["ops"][name]["roofline"] = {"algorithmic_bytes": round(b), "achieved_GBps": round(g, 1),
                                                 "frac": round(g / HBM_PEAK_GBPS, 4)}
        # the answers, against numpy
        if fmt == "buff":
            bad = verify_plain(ctx, ch, x, [pred10], [-1], words, out, buff_amd)
        else:
            bad = verify_major(ch, x, [("range", pred10), ("equal", in_major), ("equal", outlier)], words, out,
                               buff_amd)
        if bad:
            log("bench: WRONG RESULTS (%s, outlier-heavy): %s" % (fmt, bad))
            sys.exit(1)
        fres["verified"] = True
        ch.free()
        res[fmt] = fres
    for b in (xb, words, out):
        b.free()
    return res


def bench_residue(ctx, args, buff_amd, datagen):
    """RAPG residue format (the p_scan driver's codec, pscan/main.rs): a column of
    95% N(100, 0.5) and 1% 100 + 10^U[3,7] spikes, 3 decimals, seed 11, so the
    top byte columns are mostly zero. byte_residue_encode on the GPU, then
    byte_residue_decode and byte_residue_sum at full precision and at 0..3
    decimals, each timed with the library's HIP events (median of 5). The answers
    are checked against numpy (full decode = the fixed grid, sums against the
    exact integer sums) and the oracle's restatement is timed on one core on a
    sample."""
    n = args.residue_rows
    scale, dlen = 1000, 11
    x = datagen.positive_spikes(n, 3, 11, 0.01)
    xb = ctx.alloc(x.nbytes)
    xb.upload(x)
    out = ctx.alloc(n * 8)
    reps = 5
    res = {"workload": "residue (RAPG) column, %d float64 rows: 95%% N(100,0.5), 1%% 100 + 10^U[3,7], 3 decimals"
           % n, "rows": n}

    def timed(fn, kernels):
        per = []
        for _ in range(reps):
            ctx.synchronize()
            ctx.prof_reset()
            ctx.prof_enable(True)
            fn()
            ctx.synchronize()
            ctx.prof_enable(False)
            per.append(sum(ctx.prof_query(k)[1] for k in kernels) * 1e3)
        return float(np.median(per))

    enc_k = ("k_minmax", "k_major_encode", "k_major_scan", "k_major_place", "k_encode")
    us = timed(lambda: ctx.encode_device(xb.ptr, n, scale, flags=buff_amd.BUFF_RESIDUE).free(), enc_k)
    ch = ctx.encode_device(xb.ptr, n, scale, flags=buff_amd.BUFF_RESIDUE)
    info = ch.info()
    n_out = [ch.sparse_col(j)[1] for j in range(info.sparse_cols)]
    res.update({"F": info.width, "outlier_cols": info.sparse_cols, "outliers": n_out, "planes": info.planes,
                "tail_bits": info.tail_bits, "stream_bytes": None, "device_bytes": info.device_bytes})
    res["encode"] = {"device_us": round(us, 2), "GBps_f64_input": round(8.0 * n / (us * 1e-6) / 1e9, 1)}
    dense_b = n * (info.planes + info.tail_bits / 8.0)
    sparse_b = float(sum(n_out))
    ops = {}
    for p in (-1, 0, 1, 2, 3):
        us = timed(lambda: ch.materialize_device(out.ptr, p), ("k_major_decode", "k_materialize"))
        ops["decode" if p < 0 else "decode_p%d" % p] = {"device_us": round(us, 2),
                                                       "GBps_f64_output": round(8.0 * n / (us * 1e-6) / 1e9, 1)}
        us = timed((lambda: ch.sum()) if p < 0 else (lambda: ch.sum_precision(p)),
                   ("k_residue_sum", "k_tail_fields"))
        e = {"device_us": round(us, 2), "GBps_f64_input": round(8.0 * n / (us * 1e-6) / 1e9, 1)}
        if p < 0:   # every outlier byte, every plane and the tail
            b = sparse_b + dense_b + info.sparse_cols * ((n + 1023) // 1024) * 4.0
            e.update({"algorithmic_bytes": round(b), "achieved_GBps": round(b / (us * 1e-6) / 1e9, 1),
                      "frac": round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4)})
        ops["sum" if p < 0 else "sum_p%d" % p] = e
    res["ops"] = ops
    # the answers, against numpy
    f = np.trunc(np.ldexp(x, dlen)).astype(np.int64)
    ch.materialize_device(out.ptr, -1)
    got = out.download(np.float64, n)
    bad = []
    if not np.array_equal(got, f / 2.0 ** dlen):
        bad.append("full decode")
    _, exact, _ = ch.sum()
    tot = u64_sum((f - f.min()).astype(np.uint64)) + n * int(f.min())
    if abs(exact - tot / 2.0 ** dlen) > 1e-12 * abs(tot / 2.0 ** dlen):
        bad.append("exact sum")
    if bad:
        log("bench: WRONG RESULTS (residue): %s" % bad)
        sys.exit(1)
    res["verified"] = True
    res["stream_bytes"] = len(ch.serialize())
    ch.free()
    for b in (xb, out):
        b.free()
    if not args.no_cpu:   # the oracle on one core, on a 4M-row sample of the same column
        from oracle import refcpu

        refcpu.build()
        m = 1 << 22
        xs = x[:m]
        cpu = {"rows": m, "cores": 1, "kind": "port"}
        t = time.perf_counter()
        st, _ = refcpu.residue_encode(xs, scale)
        cpu["encode_GBps"] = round(8.0 * m / (time.perf_counter() - t) / 1e9, 4)
        t = time.perf_counter()
        refcpu.residue_decode(st)
        cpu["decode_GBps"] = round(8.0 * m / (time.perf_counter() - t) / 1e9, 4)
        t = time.perf_counter()
        refcpu.residue_sum(st)
        cpu["sum_GBps"] = round(8.0 * m / (time.perf_counter() - t) / 1e9, 4)
        res["cpu_port"] = cpu
    return res


def bench_config4(ctx, args, buff_amd):
    """Batched form: chunks with lengths log-uniform in [1K, 4M], each a seeded
    random walk (x0 = 0, N(0,1) steps, chunk seed 5 + i) rounded to a per-chunk
    precision uniform in {1..6}, generated on the host and uploaded before any
    timed region. One batch encode (a min/max
    launch and an encode launch over every tile of every chunk), then a batch of
    mixed queries (1/4 each equality, range, sum, max; predicates are values of a
    random row of the queried chunk) as one captured graph launch, with the shapes
    the batch autotuner picks, against the same queries issued one call per query."""
    from concurrent.futures import ThreadPoolExecutor

    from buff_amd import datagen

    nchunks, nq = args.config4_chunks, args.config4_queries
    rng = np.random.default_rng(5)
    lens = np.floor(np.exp(rng.uniform(np.log(1024), np.log(4 * 1024 * 1024), nchunks))).astype(np.int64)
    precs = rng.integers(1, 7, nchunks)
    row_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    total = int(row_off[-1])
    xh = np.empty(total, dtype=np.float64)
    t0 = time.time()

    def gen(i):
        a, b = int(row_off[i]), int(row_off[i + 1])
        xh[a:b] = datagen.random_walk(b - a, int(precs[i]), 5 + i)

    with ThreadPoolExecutor(16) as ex:
        list(ex.map(gen, range(nchunks)))
    xd = ctx.alloc(total * 8)
    xd.upload(xh)
    log("bench: batched data %d chunks, %d rows (%.1f GB) in %.1fs" % (nchunks, total, total * 8 / 1e9,
                                                                      time.time() - t0))
    scales = (10 ** precs).astype(np.uint64)
    res = {"workload": "%d chunks, lengths log-uniform in [1024, 4194304], seeded random walks at 1..6 decimals" %
                       nchunks, "chunks": nchunks, "rows": total}
    # ---- one-launch batch encode
    times = []
    batch = None
    for rep in range(3):
        if batch is not None:
            batch.free()
        ctx.synchronize()
        t = time.perf_counter()
        batch = ctx.batch_encode_device(xd.ptr, row_off, scales)
        ctx.synchronize()
        times.append(time.perf_counter() - t)
    enc = min(times)
    res["encode"] = {"seconds": round(enc, 5), "GBps_f64_input": round(total * 8 / enc / 1e9, 1)}
    # ---- the query batch
    qrng = np.random.default_rng(6)
    qchunk = qrng.integers(0, nchunks, nq)
    qrow = (row_off[qchunk] + (qrng.random(nq) * lens[qchunk]).astype(np.uint64)).astype(np.int64)
    preds = xh[qrow]
    ops = np.array([buff_amd._native.BUFF_Q_EQ, buff_amd._native.BUFF_Q_GT, buff_amd._native.BUFF_Q_SUM,
                    buff_amd._native.BUFF_Q_MAX])[qrng.integers(0, 4, nq)]
    qs = [(int(o), int(c), float(p), -1, 0) for o, c, p in zip(ops, qchunk, preds)]
    scanned = float(sum(int(lens[c]) for c in qchunk)) * 8.0
    tuned = batch.tune(qs[: min(nq, 2048)])
    res["autotuner"] = {"grid": tuned[0], "tiles_per_item": tuned[1], "first_filter_planes": tuned[2]}
    # the descriptors in buff_query layout, as a native caller holds them (their
    # conversion from Python tuples is a cost of this binding, not of the call)
    pq = buff_amd._native.prepare_queries(qs)
    bt = []
    for rep in range(5):
        ctx.synchronize()
        t = time.perf_counter()
        out = batch.query(pq)
        bt.append(time.perf_counter() - t)
    bt.sort()
    assert all(r.status == 0 for r in out)
    # the batch's answers (the last run's, still in its arena) against numpy on a sample of the queries
    t = time.time()
    sample = np.random.default_rng(7).choice(nq, min(nq, 2048), replace=False)
    bms = {int(j): batch.bitmap(out[j], int(lens[qs[j][1]])) for j in sample
           if qs[j][0] in (buff_amd._native.BUFF_Q_GT, buff_amd._native.BUFF_Q_EQ)}

    def check(j):
        from fractions import Fraction
        N = buff_amd._native
        op, c, pred, _, _ = qs[j]
        r = out[j]
        info = batch.chunks[c].info()
        x = xh[int(row_off[c]):int(row_off[c + 1])]
        off = stored_offsets(x, info)
        n, F, dlen, mn = info.n, info.width, info.dlen, info.base
        if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
            fp = int(_fixed(np.array([pred]), dlen)[0])
            if fp < mn:
                m = np.full(n, op == N.BUFF_Q_GT)
            else:
                tt = fp - mn if F < 8 else (fp - mn) & ((1 << F) - 1)
                m = off > np.uint64(tt) if op == N.BUFF_Q_GT else off == np.uint64(tt)
            if op == N.BUFF_Q_EQ and F >= 8 and F % 8 == 0:
                m = np.zeros(n, bool)   # the equality answer stays in rb1 (buff_simd.rs:4987-5091)
            w = bms[j]
            return r.count == int(m.sum()) and np.array_equal(w, _words64(m).view(np.uint32)[: len(w)])
        if op == N.BUFF_Q_SUM:
            so = u64_sum(off)
            return r.exact == float(Fraction(n * mn + so, 1 << dlen))
        mo = int(off.max())
        return r.max_off == mo and r.value == float(mo + mn) / 2.0 ** dlen

    with ThreadPoolExecutor(16) as ex:
        ok = list(ex.map(check, [int(j) for j in sample]))
    if not all(ok):
        log("bench: WRONG RESULTS (batched): queries %s" % [int(j) for j, v in zip(sample, ok) if not v][:20])
        sys.exit(1)
    res["verified_queries"] = len(sample)
    log("bench: %d batched answers verified against numpy in %.1fs" % (len(sample), time.time() - t))
    res["batched"] = {"queries": nq, "seconds": round(bt[len(bt) // 2], 5), "best_seconds": round(bt[0], 5),
                      "queries_per_s": round(nq / bt[len(bt) // 2], 1),
                      "GBps_f64_scanned": round(scanned / bt[len(bt) // 2] / 1e9, 1)}
    # the graph's device time (events around the launch) against the bytes the
    # queries must read: filters the top plane and write the bitmap, sums every
    # plane and tail bit, maxima the top plane (refinement and ties not counted)
    infos = {}
    alg = 0.0
    for (op, c, _, _, _) in qs:
        if c not in infos:
            infos[c] = batch.chunks[c].info()
        inf = infos[c]
        P, r, n_c = inf.planes, inf.tail_bits, inf.n
        top = float(((n_c + 1023) // 1024) * 1024) if P else n_c * r / 8.0
        if op in (buff_amd._native.BUFF_Q_GT, buff_amd._native.BUFF_Q_EQ):
            alg += top + (n_c + 63) // 64 * 8
        elif op == buff_amd._native.BUFF_Q_SUM:
            alg += n_c * (P + r / 8.0)
        else:
            alg += top
    gt = []
    for rep in range(3):
        ctx.synchronize()
        ctx.prof_reset()
        ctx.prof_enable(True)
        batch.query(pq)
        ctx.synchronize()
        ctx.prof_enable(False)
        launches, ms = ctx.prof_query("k_batch_graph")
        if launches:
            gt.append(ms / launches)
    if gt:
        dev = sorted(gt)[len(gt) // 2] * 1e-3
        res["batched"].update({"graph_device_s": round(dev, 6), "algorithmic_bytes": round(alg),
                               "hbm_GBps": round(alg / dev / 1e9, 1),
                               "hbm_frac": round(alg / dev / 1e9 / HBM_PEAK_GBPS, 4),
                               "host_share": round(1.0 - dev / bt[len(bt) // 2], 3)})
    # ---- the same queries, one call each
    words = ctx.alloc(((int(lens.max()) + 63) // 64) * 8)
    ctx.synchronize()
    t = time.perf_counter()
    for (op, c, p, _, _) in qs:
        ch = batch.chunks[c]
        if op == buff_amd._native.BUFF_Q_GT:
            ch.filter_device("range", p, words.ptr)
        elif op == buff_amd._native.BUFF_Q_EQ:
            ch.filter_device("equal", p, words.ptr)
        elif op == buff_amd._native.BUFF_Q_SUM:
            ch.sum()
        else:
            ch.max()
    ctx.synchronize()
    single = time.perf_counter() - t
    res["one_call_per_query_python"] = {"seconds": round(single, 5), "queries_per_s": round(nq / single, 1),
                                        "note": "ctypes calls from a Python loop"}
    # the same calls from a native loop (buff_trace_per_call, every arrival at 0:
    # each query is its single-chunk call, a filter followed by a stream
    # synchronization; a sum or max is done when its call returns)
    ctx.trace_per_call(batch.chunks, qs[:16], np.zeros(16))   # first launches of every kernel shape
    ctx.synchronize()
    t = time.perf_counter()
    _, lat = ctx.trace_per_call(batch.chunks, pq, np.zeros(len(qs)))
    native = time.perf_counter() - t
    res["one_call_per_query"] = {"seconds": round(native, 5), "queries_per_s": round(nq / native, 1),
                                 "last_completion_s": round(float(lat[-1]) * 1e-6, 5),
                                 "note": "native loop: one single-chunk call per query (a filter then a stream sync)"}
    res["batched_speedup"] = round(native / bt[len(bt) // 2], 2)
    res["batched_speedup_vs_python_loop"] = round(single / bt[len(bt) // 2], 2)
    words.free()
    batch.free()
    xd.free()
    return res


def bench_config5(ctx_unused, args, buff_amd):
    """Served form: outlier-heavy chunks (95% N(100,0.5), 5% +-10^U[6,9], 3
    decimals) with lengths log-uniform in [2^20, 2^26], encoded in the buff-major
    layout (buff_encode_majority, the layout the outlier-heavy column defines) and kept resident until
    the compressed bytes reach the target (the chunk data repeats a pool of 64
    distinct generated columns), in a context of its own with a stated HBM
    budget. A Poisson request stream (60% range filters at a random predicate,
    25% sums, 15% materializations at 1-3 decimals) is replayed open-loop by
    native clients at fixed rates; p50/p99 latency and throughput are reported
    for continuous batching (the server), for the server at one request per
    batch, and for plain single-chunk calls, one per request (the per-query-launch
    baseline). Outputs of filters and materializations go to the server's arena
    (admission by memory) and are returned as their requests complete."""
    from concurrent.futures import ThreadPoolExecutor
    from fractions import Fraction

    from buff_amd import datagen
    from buff_amd import _native as N

    target = args.config5_gb * 1e9
    budget = int(target * 1.5)
    ctx = buff_amd.Context(0, hbm_budget=budget)
    rng = np.random.default_rng(17)
    K = 64
    plens = np.floor(np.exp(rng.uniform(np.log(1 << 20), np.log(1 << 26), K))).astype(np.int64)
    poff = np.concatenate([[0], np.cumsum(plens)]).astype(np.int64)
    xh = np.empty(int(poff[-1]), dtype=np.float64)

    def gen(i):
        xh[poff[i]:poff[i + 1]] = datagen.outlier_heavy(int(plens[i]), 3, 1000 + i)

    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        list(ex.map(gen, range(K)))
    xd = ctx.alloc(xh.nbytes)
    xd.upload(xh)
    chunks, lens = [], []
    resident = 0
    while resident < target:
        i = len(chunks) % K
        try:
            ch = ctx.encode_device(xd.ptr + 8 * int(poff[i]), int(plens[i]), 10 ** 3, flags=buff_amd.BUFF_MAJORITY)
        except N.OutOfMemory:
            break   # the HBM budget admits no more chunks
        chunks.append(ch)
        lens.append(int(plens[i]))
        resident += ch.info().device_bytes
    ctx.synchronize()
    log("bench: served store %d chunks, %.1f GB resident, %d rows, built in %.1fs" % (
        len(chunks), resident / 1e9, sum(lens), time.time() - t0))
    arena_bytes = 8 << 30
    res = {"workload": "outlier-heavy chunks, lengths log-uniform in [2^20, 2^26], 3 decimals, buff-major layout",
           "chunks": len(chunks), "resident_bytes": int(resident), "rows": int(sum(lens)),
           "mix": "60% range filter, 25% sum, 15% materialize at 1-3 decimals (semantic)",
           "hbm_budget_bytes": budget, "server_arena_bytes": arena_bytes,
           "server_slots": args.config5_slots, "server_max_batch": args.config5_max_batch}
    nreq = args.config5_requests

    # filter counts and exact sums of every pool column, for checking the served answers
    qpreds = (99.5, 100.0, 100.5, 1e6)

    def pool_answers(i):
        # the chunks store every offset bit (the buff-major encoder's min/max is exact)
        info = chunks[i].info()
        off = (_fixed(xh[poff[i]:poff[i + 1]], info.dlen) - info.base).astype(np.uint64)
        majors = [chunks[i].sparse_col(j)[0] for j in range(info.sparse_cols)]
        counts = {}
        for pred in qpreds:
            counts[pred] = major_range_count(off, info, majors, int(_fixed(np.array([pred]), info.dlen)[0]))
        return counts, float(Fraction(info.n * info.base + u64_sum(off), 1 << info.dlen))

    with ThreadPoolExecutor(16) as ex:
        answers = list(ex.map(pool_answers, range(min(K, len(chunks)))))

    def trace(seed):
        r = np.random.default_rng(seed)
        qs, rows = [], 0
        for _ in range(nreq):
            k = int(r.integers(0, len(chunks)))
            u = r.random()
            if u < 0.6:
                qs.append((N.BUFF_Q_GT, k, float(r.choice(qpreds)), -1, 0))
            elif u < 0.85:
                qs.append((N.BUFF_Q_SUM, k, 0.0, -1, 0))
            else:
                qs.append((N.BUFF_Q_MAT, k, 0.0, int(r.integers(1, 4)), N.BUFF_SEMANTIC))
            rows += lens[k]
        return qs, rows

    def check(qs, out, mode, counts=True):
        wrong = [i for i, (q, o) in enumerate(zip(qs, out))
                 if o.status != 0 or (counts and q[0] == N.BUFF_Q_GT and o.count != answers[q[1] % K][0][q[2]]) or
                 (q[0] == N.BUFF_Q_SUM and o.exact != answers[q[1] % K][1])]
        if wrong:
            log("bench: WRONG RESULTS (served, %s): requests %s" % (mode, wrong[:20]))
            sys.exit(1)

    def serve(mode, qps, seed):
        qs, rows = trace(seed)
        arrival = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / qps, nreq)) if qps else np.zeros(nreq)
        extra = {}
        pq = buff_amd._native.prepare_queries(qs)   # buff_query descriptors, built before the clock starts
        if mode == "plain_per_call":
            ctx.trace_per_call(chunks, qs[:8], np.arange(8) * 1e-3)   # first launches of every kernel shape
            t = time.perf_counter()
            out, lat = ctx.trace_per_call(chunks, pq, arrival)
            wall = time.perf_counter() - t
            check(qs, out, mode, counts=False)
        else:
            mb, sl = (args.config5_max_batch, args.config5_slots) if mode == "continuous_batching" else (1, 1)
            srv = buff_amd.Server(ctx, max_batch=mb, slots=sl, arena_bytes=arena_bytes)
            try:
                for ch in chunks:
                    srv.add_chunk(ch)
                # views upload and first launches: requests 1 ms apart, so each is a batch of one and the largest
                # batch reported below is the timed trace's
                srv.run_trace(qs[:8], np.arange(8) * 1e-3)
                b0 = srv.stats()[0]
                t = time.perf_counter()
                out, lat = srv.run_trace(pq, arrival)
                wall = time.perf_counter() - t
                batches, _, largest = srv.stats()
                extra = {"batches": int(batches - b0), "largest_batch": int(largest),
                         "arena_peak_bytes": int(srv.arena_stats()[1])}
            finally:
                srv.destroy()
            check(qs, out, mode)
        d = {"offered_qps": qps or None, "achieved_qps": round(nreq / wall, 1),
             "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
             "GBps_f64_scanned": round(rows * 8 / wall / 1e9, 1)}
        d.update(extra)
        return d

    modes = ("continuous_batching", "one_launch_per_request", "plain_per_call")
    res["saturated"] = {m: serve(m, 0, 50) for m in modes}
    sat = {m: res["saturated"][m]["achieved_qps"] for m in modes}
    res["speedup_saturated"] = {"vs_plain_per_call": round(sat[modes[0]] / sat["plain_per_call"], 2),
                                "vs_one_launch_per_request": round(sat[modes[0]] / sat["one_launch_per_request"],
                                                                   2)}
    sweep = []
    for frac in (0.25, 0.5, 0.8):
        qps = max(100.0, sat[modes[0]] * frac)
        sweep.append(dict({"load": frac}, **{m: serve(m, qps, 60 + int(frac * 100)) for m in modes}))
    res["poisson_sweep"] = sweep
    # throughput at a stated p99: the highest Poisson rate (bisection, 7 probes per bound) whose p99 latency stays
    # within the bound and which keeps up with its offered rate; a looser bound starts from the tighter bound's
    # best rate (which also meets it), so its throughput is never reported lower
    slo = []
    best_prev = {m: None for m in modes}
    for bound_us in (1000.0, 5000.0):
        row = {"p99_bound_us": bound_us}
        for m in modes:
            best = best_prev[m]
            lo_rate = best["offered_qps"] if best else 0.0
            hi_rate = sat[m] * 1.05
            for it in range(7):
                rate = (lo_rate + hi_rate) / 2
                if rate <= lo_rate:
                    break
                r = serve(m, rate, 200 + it + int(bound_us))
                if r["p99_us"] <= bound_us and r["achieved_qps"] >= 0.97 * rate:
                    lo_rate, best = rate, r
                else:
                    hi_rate = rate
            row[m] = best
            best_prev[m] = best
        b = row[modes[0]]
        row["ratio_vs_plain_per_call"] = (round(b["achieved_qps"] / row["plain_per_call"]["achieved_qps"], 2)
                                          if b and row["plain_per_call"] else None)
        row["ratio_vs_one_launch_per_request"] = (
            round(b["achieved_qps"] / row["one_launch_per_request"]["achieved_qps"], 2)
            if b and row["one_launch_per_request"] else None)
        slo.append(row)
    res["throughput_at_p99"] = slo
    res["hbm_used_bytes"] = int(ctx.hbm_used())
    for ch in chunks:
        ch.free()
    xd.free()
    ctx.destroy()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one MI355X, and this build has no multi-GPU path "
              "(chunks shard independently; the design is in DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import buff_amd
    from buff_amd import datagen

    n = args.rows
    t0 = time.time()
    log("bench: generating 2 x %d rows" % n)
    x_walk = datagen.random_walk(n, 4, 2)
    x_uni = datagen.uniform(n, -1000.0, 1000.0, 2, 3)
    cols = [Column("walk", x_walk, 10 ** 4, [-1, 2]), Column("uniform", x_uni, 10 ** 2, [-1])]
    log("bench: data ready in %.1fs" % (time.time() - t0))

    ctx = buff_amd.Context(0)
    dx = []
    for c, x in zip(cols, (x_walk, x_uni)):
        b = ctx.alloc(x.nbytes)
        b.upload(x)
        dx.append(b)
    words = ctx.alloc(((n + 63) // 64) * 8)
    out = ctx.alloc(n * 8)
    flush = ctx.alloc(1 << 30)

    def step(record=None):
        chunks = []
        for c, b in zip(cols, dx):
            chunks.append(ctx.encode_device(b.ptr, c.n, c.scale))
        ops = 0
        for sel in (0.001, 0.10, 0.50):
            for c, ch in zip(cols, chunks):
                ch.filter_device("range", c.preds[sel], words.ptr)
                ops += 1
        for ch in chunks:
            ch.sum()
            ch.max()
            ops += 2
        for c, ch in zip(cols, chunks):
            for p in c.mat_precisions:
                ch.materialize_device(out.ptr, p, flags=buff_amd.BUFF_SEMANTIC)
                ops += 1
        ctx.synchronize()
        if record is not None:
            record.extend(chunks)
        else:
            for ch in chunks:
                ch.free()
        return ops + len(chunks)

    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    # launch-shape autotuner: each column's chunk, at its 10% predicate (kept per size class and byte planes)
    tuned = {}
    if not args.no_tune:
        for c, b in zip(cols, dx):
            ch = ctx.encode_device(b.ptr, c.n, c.scale)
            tuned[c.name] = ch.tune(c.preds[0.10])
            ch.free()
        log("bench: tuned launch shapes %s" % json.dumps(tuned))
    log("bench: timing %d steps" % args.steps)
    ctx.synchronize()
    t = time.perf_counter()
    ops = 0
    for i in range(args.steps):
        ops = step()
        log("bench: step %d done" % (i + 1))
    ctx.synchronize()
    dt = time.perf_counter() - t
    # The same K steps again with per-kernel HIP events (each profiled kernel
    # launch carries start/stop events that time its execution): launching
    # with events costs host time between operations, so it stays out of the
    # timed steps above.
    ctx.prof_reset()
    ctx.prof_enable(True)
    for i in range(args.steps):
        step()
    ctx.synchronize()
    ctx.prof_enable(False)
    ms_per_step = dt * 1e3 / args.steps
    logical_bytes = 8.0 * n * ops
    value = logical_bytes / (dt / args.steps) / 1e9

    # ---- the step's answers, checked against numpy (no line is printed if any is wrong)
    keep = []
    step(record=keep)
    t = time.time()
    bad = []
    for c, ch, x in zip(cols, keep, (x_walk, x_uni)):
        bad += ["%s: %s" % (c.name, b) for b in verify_plain(ctx, ch, x, [c.preds[s_] for s_ in (0.001, 0.10, 0.50)],
                                                             c.mat_precisions, words, out, buff_amd)]
    if bad:
        log("bench: WRONG RESULTS: %s" % bad)
        sys.exit(1)
    log("bench: step results verified against numpy in %.1fs" % (time.time() - t))

    # ---- per-kernel accounting over the timed region ----------------------------
    infos = [ch.info() for ch in keep]
    stats_dev = ctx.alloc(9 * 8)
    filt = {}
    for c, ch, info in zip(cols, keep, infos):
        for sel in (0.001, 0.10, 0.50):
            stats_dev.fill(0)
            ch.filter_device("range", c.preds[sel], words.ptr, stats_dev=stats_dev.ptr)
            filt[(c.name, sel)] = algorithmic_bytes(info, "k_filter_gt", stats_dev.download(np.uint64, 9))
    mat = 0.0
    for c, info in zip(cols, infos):
        for p in c.mat_precisions:
            mat += algorithmic_bytes(info, "k_materialize", precision=precision_R(info, p))
    per_step = {
        "k_minmax": sum(algorithmic_bytes(i, "k_minmax") for i in infos),
        "k_encode": sum(algorithmic_bytes(i, "k_encode") for i in infos),
        "k_filter_gt": sum(filt.values()),
        "k_sum": sum(algorithmic_bytes(i, "k_sum") for i in infos),
        "k_max": sum(algorithmic_bytes(i, "k_max") for i in infos),
        "k_materialize": mat,
    }
    kernels = {}
    for k, bytes_per_step in per_step.items():
        launches, total_ms = ctx.prof_query(k)
        if not launches:
            continue
        avg_ms = total_ms / launches
        bytes_per_launch = bytes_per_step / (launches / args.steps)
        gbps = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        kernels[k] = {"launches": launches, "avg_us": round(avg_ms * 1e3, 2),
                      "share_of_step": round(total_ms / args.steps / ms_per_step, 4),
                      "algorithmic_bytes_per_launch": round(bytes_per_launch),
                      "achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}
    dominant = max(kernels, key=lambda k: kernels[k]["share_of_step"])

    # ---- cold-cache pass: evict the memory-side cache before every launch ------
    cold_bytes = {"k_filter_gt": [filt[(c.name, 0.10)] for c in cols],
                  "k_sum": [algorithmic_bytes(i, "k_sum") for i in infos],
                  "k_max": [algorithmic_bytes(i, "k_max") for i in infos],
                  "k_materialize": [algorithmic_bytes(i, "k_materialize") for i in infos]}
    cold = {}
    trials = 3
    line_bytes = {}
    for c, info, x in zip(cols, infos, (x_walk, x_uni)):
        off = stored_offsets(x, info)
        for sel in (0.001, 0.10, 0.50):
            line_bytes[(c.name, sel)] = filter_line_bytes(off, info, int(_fixed(np.array([c.preds[sel]]),
                                                                                info.dlen)[0]))
        del off
    # each range filter alone, cold: per column and selectivity
    filters = {}
    for c, ch in zip(cols, keep):
        for sel in (0.001, 0.10, 0.50):
            ctx.prof_reset()
            for trial in range(trials):
                flush.touch()
                ctx.synchronize()
                ctx.prof_enable(True)
                ch.filter_device("range", c.preds[sel], words.ptr)
                ctx.synchronize()
                ctx.prof_enable(False)
            launches, total_ms = ctx.prof_query("k_filter_gt")
            if launches:
                us = total_ms / launches * 1e3
                gbps = filt[(c.name, sel)] / (us * 1e-6) / 1e9
                lb = line_bytes[(c.name, sel)]
                filters["%s@%g%%" % (c.name, sel * 100)] = {"avg_us": round(us, 2), "achieved_GBps": round(gbps, 1),
                                                            "frac": round(gbps / HBM_PEAK_GBPS, 4),
                                                            "algorithmic_bytes": round(filt[(c.name, sel)]),
                                                            "line_granular_bytes": round(lb),
                                                            "frac_of_line_granular_roofline":
                                                                round(lb / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4)}
    ctx.prof_reset()
    for trial in range(trials):
        for c, ch in zip(cols, keep):
            for fn in (lambda: ch.filter_device("range", c.preds[0.10], words.ptr), lambda: ch.sum(),
                       lambda: ch.max(), lambda: ch.materialize_device(out.ptr, -1)):
                flush.touch()
                ctx.synchronize()
                ctx.prof_enable(True)
                fn()
                ctx.synchronize()
                ctx.prof_enable(False)
    for k, per_col in cold_bytes.items():
        launches, total_ms = ctx.prof_query(k)
        if launches:
            b = sum(per_col) / len(per_col)
            gbps = b / ((total_ms / launches) * 1e-3) / 1e9
            cold[k] = {"launches": launches, "avg_us": round(total_ms / launches * 1e3, 2),
                       "achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
                       "note": "10% filter" if k == "k_filter_gt" else ("full precision" if k == "k_materialize"
                                                                       else "")}
    # the device's streaming read + write rate beside the kernels that read and
    # write: a plain non-temporal copy of the walk column's float64 (n x 8
    # bytes) into the materialize output, warm and after a cache eviction
    def ceiling(kernel, nbytes_moved, launch, note):
        sec = {"bytes_per_launch": nbytes_moved, "note": note}
        for label, evict in (("warm", False), ("cold", True)):
            ctx.prof_reset()
            for trial in range(trials + 1):   # the first launch is not timed
                if evict:
                    flush.touch()
                ctx.synchronize()
                ctx.prof_enable(trial > 0)
                launch()
                ctx.synchronize()
                ctx.prof_enable(False)
            launches, total_ms = ctx.prof_query(kernel)
            if launches:
                us = total_ms / launches * 1e3
                gbps = float(nbytes_moved) / (us * 1e-6) / 1e9
                sec[label] = {"avg_us": round(us, 2), "achieved_GBps": round(gbps, 1),
                              "frac": round(gbps / HBM_PEAK_GBPS, 4)}
        return sec

    stream_copy = ceiling("k_copy_bytes", 2 * n * 8, lambda: out.copy_from(dx[0], n * 8),
                          "k_copy_bytes: n x 8 bytes read + n x 8 bytes written")
    # the streaming read of one plane's worth of bytes (n: a top byte plane),
    # the floor of the plane-reading kernels at this size, and of the float64
    # input (k_minmax's read)
    stream_read = ceiling("k_read_bytes", n, lambda: dx[0].stream_read(n),
                          "k_read_bytes: n bytes read (one byte plane), non-temporal")
    stream_read_f64 = ceiling("k_read_bytes", n * 8, lambda: dx[0].stream_read(n * 8),
                              "k_read_bytes: n x 8 bytes read (the float64 input), non-temporal")
    # the streaming write of n float64 (materialize's output)
    stream_write = ceiling("k_fill_bytes", n * 8, lambda: out.fill(0), "k_fill_bytes: n x 8 bytes written")
    filter_project = bench_filter_project(ctx, cols[0], keep[0], infos[0], words, out, flush, buff_amd)
    window_max = bench_window_max(ctx, cols[0], keep[0], infos[0], words, out, flush, buff_amd)
    slice_sec = bench_slice(ctx, cols[0], dx[0], words, out, flush, buff_amd)
    ingest = bench_ingest(ctx, flush)
    for ch in keep:
        ch.free()

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "value_is": "logical GB/s: 8 bytes of float64 input per row per operation over the step; hbm_GBps is "
                    "the step's algorithmic HBM bytes over the same time",
        "hbm_GBps": round(sum(per_step.values()) / (ms_per_step * 1e-3) / 1e9, 1),
        "hbm_frac": round(sum(per_step.values()) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64->int64 (byte planes, u64 sums)",
        "data": "synthetic (seeded random walk 4 decimals + uniform[-1000,1000] 2 decimals)",
        "config": {"workload": "plain BUFF, 2 columns x %d float64 rows (walk p=4, uniform p=2): encode, "
                               "range filters at 0.1/10/50%%, sum, max, materialize" % n,
                   "rows_per_column": n, "columns": 2, "ops_per_step": ops, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": kernels[dominant]["achieved_GBps"],
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": kernels[dominant]["frac"],
                     "traffic": pmc_traffic(dominant, n), "traffic_source": TRAFFIC_FILE},
        "kernels": kernels,
        "launch_shapes": tuned,
        "cold_cache": cold,
        "cold_filters": filters,
        "stream_copy": stream_copy,
        "stream_read_plane": stream_read,
        "stream_read_f64": stream_read_f64,
        "stream_write": stream_write,
        "filter_project": filter_project,
        "window_max": window_max,
        "slice": slice_sec,
        "ingest": ingest,
    }

    if not args.no_config3:
        for b in (flush, words, out):
            b.free()
        for b in dx:
            b.free()
        log("bench: outlier-heavy section (%d rows)" % args.config3_rows)
        result["config3"] = bench_config3(ctx, args, buff_amd, datagen)
    if not args.no_residue:
        log("bench: residue section (%d rows)" % args.residue_rows)
        result["residue"] = bench_residue(ctx, args, buff_amd, datagen)
    if not args.no_config4:
        log("bench: batched section (%d chunks, %d queries)" % (args.config4_chunks, args.config4_queries))
        result["config4"] = bench_config4(ctx, args, buff_amd)
    if not args.no_config5:
        log("bench: served section (%.0f GB resident)" % args.config5_gb)
        result["config5"] = bench_config5(ctx, args, buff_amd)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_rows)
        result["slice"]["cpu_port"] = cpu_slice(1 << 24)
    print(json.dumps(result), flush=True)


def cpu_slice(rows):
    """The oracle's restatement of the byte-slice path on one core, on a bounded
    sample of the walk column: buff_slice_encode, buff_slice_range_filter at
    10%, buff_slice_decode. GB/s of float64 input per operation."""
    from buff_amd import datagen
    from oracle import refcpu

    refcpu.build()
    x = datagen.random_walk(rows, 4, 2)
    res = {"rows": rows, "cores": 1, "kind": "port"}
    t = time.perf_counter()
    s = refcpu.slice_encode(x, 10 ** 4)
    res["encode_GBps"] = round(8.0 * rows / (time.perf_counter() - t) / 1e9, 4)
    pred = float(np.quantile(x, 0.90, method="lower"))
    t = time.perf_counter()
    refcpu.slice_filter(s, 10 ** 4, "range", pred)
    res["range_10pct_GBps"] = round(8.0 * rows / (time.perf_counter() - t) / 1e9, 4)
    t = time.perf_counter()
    refcpu.slice_decode(s)
    res["decode_GBps"] = round(8.0 * rows / (time.perf_counter() - t) / 1e9, 4)
    return res


def cpu_baseline(rows):
    """The oracle's restatement of the reference CPU path on a bounded sample.

    buff-simd (AVX2) encode/decode/range filters and the scalar buff sum/max,
    single-threaded as the reference, on `rows` rows of each column.
    """
    from buff_amd import datagen
    from oracle import refcpu

    refcpu.build()
    t_all = 0.0
    logical = 0.0
    samples = [(datagen.random_walk(rows, 4, 2), 10 ** 4), (datagen.uniform(rows, -1000.0, 1000.0, 2, 3), 10 ** 2)]
    for x, scale in samples:
        t = time.perf_counter()
        s = refcpu.encode(x, scale, flip=True)
        t_all += time.perf_counter() - t
        logical += 8.0 * rows
        for sel in (0.001, 0.10, 0.50):
            pred = float(np.quantile(x, 1.0 - sel, method="lower"))
            t = time.perf_counter()
            refcpu.simd_filter(s, scale, "range", pred)
            t_all += time.perf_counter() - t
            logical += 8.0 * rows
        plain = refcpu.encode(x, scale)
        for fn in (refcpu.byte_fixed_sum, refcpu.byte_fixed_max):
            t = time.perf_counter()
            fn(plain, scale)
            t_all += time.perf_counter() - t
            logical += 8.0 * rows
        t = time.perf_counter()
        refcpu.decode(s, scale, flip=True)
        t_all += time.perf_counter() - t
        logical += 8.0 * rows
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(logical / t_all / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d rows of each column (walk p=4, uniform p=2): buff-simd encode, 3 range filters (AVX2), "
                      "decode; scalar buff sum and max; single thread on %s (nproc %d)" % (rows, cpu,
                                                                                         os.cpu_count() or 0)}


if __name__ == "__main__":
    main()
