"""Correctness at the workloads' real sizes, through the C ABI.

The oracle is too slow for a few of these sizes, so most checks here are
properties that hold at any size, computed with numpy from the input values:
the stored offsets are off = fixed - base with fixed = trunc(x * 2^dlen)
(prec_double.rs:190-205 for finite |x| < 2^(63 - dlen)), and every answer is
a function of them (the stream's planes and tail bits, filter bitmaps,
integer sums, maxima, decoded doubles). Where the oracle finishes in seconds
(buff-major sums, filters, maxima and decodes at 67,108,864 rows) it is the
checker itself.

  * two columns of 134,217,728 rows (random walk at 4 decimals, uniform
    [-1000, 1000] at 2 decimals): header and width with the `else if` max, the
    serialized stream byte for byte against planes and tail built from the
    offsets, stream -> resident -> stream, range filters at 0.1/10/50%
    selectivity and equality on a stored value (bitmaps and counts), the exact
    sum and the reference fold, max, full and reduced-precision decode;
  * 67,108,864 outlier-heavy rows in the plain and the buff-major layout: the
    buff-major stream and its queries against the oracle, the semantic answers
    against numpy;
  * 4096 chunks of 1K-4M rows at 1..6 decimals encoded in one batch, and one
    batch of 16,384 mixed queries checked against numpy on each chunk.
"""
import math
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction

import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg

pytestmark = pytest.mark.gpu

PRECISION_MAP = {1: 5, 2: 8, 3: 11, 4: 15, 5: 18, 6: 21}


def _fixed(x, dlen):
    return np.trunc(np.ldexp(x, dlen)).astype(np.int64)


def _minmax(f):
    """min and the `else if` max of byte_fixed_encode (split_double.rs:321-330):
    max{f_i : i >= 1, f_i >= min(f_0 .. f_{i-1})}."""
    mn = int(f.min())
    if len(f) < 2:
        return mn, -(1 << 63)
    pm = np.minimum.accumulate(f[:-1])
    rest = f[1:]
    cand = rest[rest >= pm]
    return mn, int(cand.max()) if len(cand) else -(1 << 63)


def _width(mn, mx):
    d = (mx - mn) & ((1 << 64) - 1)
    d = d - (1 << 64) if d >= 1 << 63 else d
    return int(math.ceil(math.log2(float(d)))) if d else 0


def _u64_sum(a):
    """Exact sum of a uint64 array (two 32-bit halves, so no partial sum wraps)."""
    a = np.asarray(a, dtype=np.uint64)
    return (int(np.sum(a >> np.uint64(32), dtype=np.uint64)) << 32) + int(np.sum(a & np.uint64(0xFFFFFFFF),
                                                                                 dtype=np.uint64))


def _exact_fixed_sum(f):
    mn = int(f.min())
    return len(f) * mn + _u64_sum((f - mn).astype(np.uint64))


def _words(mask):
    b = np.packbits(mask, bitorder="little")
    b = np.concatenate([b, np.zeros((-len(b)) % 4, np.uint8)])
    return b.view(np.uint32)[: (len(mask) + 31) // 32]


def _check_plain_chunk(ch, x, scale, preds, mat_precisions):
    p = int(round(math.log10(scale)))
    dlen = PRECISION_MAP[p]
    f = _fixed(x, dlen)
    n = len(x)
    mn, mx = _minmax(f)
    F = _width(mn, mx)
    info = ch.info()
    assert (info.base, info.n, info.dlen, info.width) == (mn, n, dlen, F)
    assert int(f.max()) == mx, "row 0 is the unique maximum: the `else if` case is tested at small sizes"
    off = (f - mn).astype(np.uint64)
    P, r = (F // 8, F % 8) if F >= 8 else (0, F)
    # the serialized stream, byte for byte
    s = ch.serialize()
    assert len(s) == 20 + P * n + (n * r + 7) // 8
    hdr = np.frombuffer(s[:20].tobytes(), dtype=np.uint32)
    assert int(hdr[0]) | (int(hdr[1]) << 32) == mn & ((1 << 64) - 1)
    assert (int(hdr[2]), int(hdr[3]), int(hdr[4])) == (n, (F - dlen) & 0xFFFFFFFF, dlen)
    for k in range(P):
        plane = (off >> np.uint64(F - 8 * (k + 1))).astype(np.uint8)
        assert np.array_equal(s[20 + k * n: 20 + (k + 1) * n], plane), ("plane", k)
    if r:
        low = (off & np.uint64((1 << r) - 1)).astype(np.uint8)
        bits = ((low[:, None] >> np.arange(r, dtype=np.uint8)[None, :]) & 1).astype(np.uint8).reshape(-1)
        assert np.array_equal(s[20 + P * n:], np.packbits(bits, bitorder="little")), "tail"
        del bits, low
    # stream -> resident -> stream
    ch2 = ch.ctx.load(s, scale)
    assert np.array_equal(ch2.serialize(), s)
    ch2.free()
    del s
    # filters: x > pred and x == pred are off > t and off == t on the grid
    for pred in preds:
        fp = int(_fixed(np.array([pred]), dlen)[0])
        assert mn <= fp <= mx
        w, c = ch.filter("range", pred)
        m = f > fp
        assert c == int(m.sum()) and np.array_equal(w, _words(m)), ("range", pred)
        if F % 8:
            w, c = ch.filter("equal", pred)
            m = f == fp
            assert c == int(m.sum()) and np.array_equal(w, _words(m)), ("equal", pred)
    # sum: exact, and the reference fold within 1e-12 of the summed magnitudes
    tot = _exact_fixed_sum(f)
    ref, exact, ps = ch.sum()
    assert exact == float(Fraction(tot, 1 << dlen))
    for k in range(P):
        assert int(ps[k]) == int(np.sum((off >> np.uint64(F - 8 * (k + 1))) & np.uint64(0xFF), dtype=np.uint64))
    mag = (n * abs(mn) + _u64_sum(off)) / 2.0 ** dlen
    want = float(Fraction(tot, 1 << dlen)) if F % 8 else float(Fraction(tot - n * mn, 1 << dlen))
    assert abs(ref - want) <= 1e-12 * mag
    # max
    v, o = ch.max()
    assert o == mx - mn and v == float(mx) / 2.0 ** dlen
    # materialize: full, and semantic reduced precision
    for prec in [-1] + list(mat_precisions):
        got = ch.materialize(prec, flags=buff_amd.BUFF_SEMANTIC)
        if prec < 0 or PRECISION_MAP[prec] > dlen:
            want_v = f.astype(np.float64) / 2.0 ** dlen
        else:
            R = (F - dlen + PRECISION_MAP[prec]) & 0xFFFFFFFF
            if R // 8 < F // 8 and R % 8:
                R = (R // 8 + 1) * 8
            R = min(R, F)
            sh = np.uint64(F - R)
            want_v = (((off >> sh) << sh).astype(np.int64) + mn).astype(np.float64) / 2.0 ** dlen
        assert np.array_equal(got.view(np.uint64), want_v.view(np.uint64)), ("materialize", prec)


@pytest.mark.timeout(240)
@pytest.mark.parametrize("col", ["walk", "uniform"])
def test_config2_column_134M(gpu_ctx, col):
    n = 1 << 27
    print("realsize: config2 column %s, %d rows" % (col, n), flush=True)
    if col == "walk":
        x, scale, mats = dg.random_walk(n, 4, 2), 10 ** 4, (2,)
    else:
        x, scale, mats = dg.uniform(n, -1000.0, 1000.0, 2, 3), 10 ** 2, ()
    preds = [float(np.quantile(x, 1.0 - sel, method="lower")) for sel in (0.001, 0.10, 0.50)]
    preds.append(float(x[n // 3]))
    xd = gpu_ctx.alloc(x.nbytes)
    xd.upload(x)
    ch = gpu_ctx.encode_device(xd.ptr, n, scale)
    xd.free()
    try:
        _check_plain_chunk(ch, x, scale, preds, mats)
    finally:
        ch.free()


@pytest.mark.timeout(240)
def test_config3_outlier_heavy_64M(gpu_ctx, oracle):
    n = 1 << 26
    scale, dlen = 10 ** 3, 11
    print("realsize: config3, %d rows, plain and buff-major" % n, flush=True)
    x = dg.outlier_heavy(n, 3, 4)
    pred10 = float(np.quantile(x, 0.90, method="lower"))
    in_major = float(x[np.argsort(np.abs(x - 100.0))[n // 3]])
    outlier = float(x[int(np.flatnonzero(np.abs(x) > 1e5)[7])])
    xd = gpu_ctx.alloc(x.nbytes)
    xd.upload(x)
    plain = gpu_ctx.encode_device(xd.ptr, n, scale)
    major = gpu_ctx.encode_device(xd.ptr, n, scale, flags=buff_amd.BUFF_MAJORITY)
    xd.free()
    try:
        # plain layout: numpy properties
        _check_plain_chunk(plain, x, scale, [pred10, in_major, outlier], (1, 2))
        # buff-major: the stream and every compat answer against the oracle
        s = major.serialize()
        assert np.array_equal(s, oracle.major_encode(x, scale)), "buff-major stream"
        for op, pred in (("range", pred10), ("equal", in_major), ("equal", outlier)):
            w_ref, c_ref = oracle.major_filter(s, scale, op, pred)[:2]
            w, c = major.filter(op, pred)
            assert c == c_ref and np.array_equal(w, w_ref), (op, pred)
        s_ref, cs_ref = oracle.major_sum(s, scale)
        r, e, cs = major.sum()
        assert r == s_ref, (r, s_ref)
        m_ref, off_ref = oracle.major_max(s, scale)[:2]
        v, o = major.max()
        assert (v, o) == (m_ref, off_ref)
        d_ref = oracle.major_decode(s, scale)
        d = major.materialize()
        assert np.array_equal(d.view(np.uint64), d_ref.view(np.uint64))
        del s, d_ref, d
        # buff-major semantic answers: numpy on the offsets
        f = _fixed(x, dlen)
        for op, pred in (("range", pred10), ("equal", in_major), ("equal", outlier)):
            fp = int(_fixed(np.array([pred]), dlen)[0])
            m = f > fp if op == "range" else f == fp
            w, c = major.filter(op, pred, flags=buff_amd.BUFF_SEMANTIC)
            assert c == int(m.sum()) and np.array_equal(w, _words(m)), ("semantic", op, pred)
        r, e, _ = major.sum(flags=buff_amd.BUFF_SEMANTIC)
        assert e == float(Fraction(_exact_fixed_sum(f), 1 << dlen))
        v, o, _ = major.max_ex(flags=buff_amd.BUFF_SEMANTIC)
        assert v == float(f.max()) / 2.0 ** dlen
        got = major.materialize(-1, flags=buff_amd.BUFF_SEMANTIC)
        assert np.array_equal(got.view(np.uint64), (f.astype(np.float64) / 2.0 ** dlen).view(np.uint64))
    finally:
        plain.free()
        major.free()


@pytest.mark.timeout(300)
def test_config4_batch_4096_chunks(gpu_ctx, oracle):
    nchunks, nq = 4096, 16384
    rng = np.random.default_rng(5)
    lens = np.floor(np.exp(rng.uniform(np.log(1024), np.log(4 * 1024 * 1024), nchunks))).astype(np.int64)
    precs = rng.integers(1, 7, nchunks)
    row_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    total = int(row_off[-1])
    print("realsize: config4, %d chunks, %d rows, %d queries" % (nchunks, total, nq), flush=True)
    xh = np.empty(total, dtype=np.float64)

    def gen(i):
        a, b = int(row_off[i]), int(row_off[i + 1])
        xh[a:b] = dg.random_walk(b - a, int(precs[i]), 5 + i)

    with ThreadPoolExecutor(16) as ex:
        list(ex.map(gen, range(nchunks)))
    xd = gpu_ctx.alloc(total * 8)
    xd.upload(xh)
    scales = (10 ** precs).astype(np.uint64)
    batch = gpu_ctx.batch_encode_device(xd.ptr, row_off, scales)
    xd.free()
    try:
        fx = [None] * nchunks

        def header(i):
            a, b = int(row_off[i]), int(row_off[i + 1])
            dlen = PRECISION_MAP[int(precs[i])]
            f = _fixed(xh[a:b], dlen)
            fx[i] = f
            mn, mx = _minmax(f)
            inf = batch.chunks[i].info()
            return (inf.base, inf.n, inf.dlen, inf.width) == (mn, b - a, dlen, _width(mn, mx))

        with ThreadPoolExecutor(16) as ex:
            ok = list(ex.map(header, range(nchunks)))
        assert all(ok), [i for i, v in enumerate(ok) if not v][:10]
        # a sample of chunks serialized against the oracle's encoder
        for i in np.random.default_rng(1).choice(nchunks, 24, replace=False):
            a, b = int(row_off[i]), int(row_off[i + 1])
            assert np.array_equal(batch.chunks[i].serialize(), oracle.encode(xh[a:b], int(scales[i]))), i
        # one batch of mixed queries (the bench's mix: 1/4 each equality, range, sum, max)
        qrng = np.random.default_rng(6)
        qchunk = qrng.integers(0, nchunks, nq)
        qrow = (row_off[qchunk] + (qrng.random(nq) * lens[qchunk]).astype(np.uint64)).astype(np.int64)
        ops = np.array([N.BUFF_Q_EQ, N.BUFF_Q_GT, N.BUFF_Q_SUM, N.BUFF_Q_MAX])[qrng.integers(0, 4, nq)]
        qs = [(int(o), int(c), float(xh[r]), -1, 0) for o, c, r in zip(ops, qchunk, qrow)]
        res = batch.query(qs)
        # filter bitmaps live in the batch arena until the next query: copy them now
        bms = {j: batch.bitmap(res[j], int(lens[qs[j][1]])) for j in range(nq)
               if qs[j][0] in (N.BUFF_Q_GT, N.BUFF_Q_EQ)}

        def check(j):
            # answers on the stored offsets: (fixed - base) mod 2^F (a chunk whose row 0 is the unique
            # maximum stores a width from the `else if` max, and its planes keep only the low F bits)
            op, c, pred, _, _ = qs[j]
            r = res[j]
            f = fx[c]
            n = len(f)
            dlen = PRECISION_MAP[int(precs[c])]
            inf = batch.chunks[c].info()
            F, mn = inf.width, inf.base
            mask = np.uint64((1 << F) - 1) if F < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
            off = (f - mn).astype(np.uint64) & mask
            if r.status != 0:
                return "status %d" % r.status
            if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
                fp = int(_fixed(np.array([pred]), dlen)[0])
                if fp < mn:
                    m = np.full(n, op == N.BUFF_Q_GT)
                else:
                    t = fp - mn if F < 8 else (fp - mn) & ((1 << F) - 1)
                    m = off > np.uint64(t) if op == N.BUFF_Q_GT else off == np.uint64(t)
                if op == N.BUFF_Q_EQ and F >= 8 and F % 8 == 0:
                    m = np.zeros(n, bool)   # the equality answer stays in rb1 (buff_simd.rs:4987-5091)
                if r.count != int(m.sum()) or not np.array_equal(bms[j], _words(m)):
                    return "filter"
            elif op == N.BUFF_Q_SUM:
                so = _u64_sum(off)
                if r.exact != float(Fraction(n * mn + so, 1 << dlen)):
                    return "exact sum"
                mag = (n * abs(mn) + so) / 2.0 ** dlen
                with_base = not (F >= 8 and F % 8 == 0)
                want = float(Fraction((n * mn if with_base else 0) + so, 1 << dlen))
                if abs(r.value - want) > 1e-12 * mag + 1e-300:
                    return "reference sum"
            else:
                mo = int(off.max())
                if r.max_off != mo or r.value != float(mo + mn) / 2.0 ** dlen:
                    return "max"
            return None

        with ThreadPoolExecutor(16) as ex:
            bad = [(j, e) for j, e in enumerate(ex.map(check, range(nq))) if e]
        assert not bad, bad[:10]
    finally:
        batch.free()
