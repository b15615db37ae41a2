"""Byte-slice chunks on the GPU through the C ABI (BuffSliceCompress,
buff_slice.rs:222-1970, and buff-simd's `_with_slice` filters, buff_simd.rs:
3736-4025) against the oracle, bit for bit: the encoded stream, loading the
oracle's stream, every filter kind (its plane cap, row coverage and panics),
the three decodes, max and argmax count, the windowed max, and the sums (the
reference's f64 running sum within 1e-12 relative, the exact sum exactly).
BUFF_SEMANTIC answers against the closed forms. At 16.8M rows, properties
against the byte_fixed chunk of the same column.
"""
from fractions import Fraction

import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg
from tests import closed_form as cf
from tests.test_slice_oracle import CASES as PIN_CASES

pytestmark = pytest.mark.gpu

CASES = [c for c in PIN_CASES] + [("walk4_100k", dg.random_walk(100_003, 4, 21), 10 ** 4)]
IDS = [c[0] for c in CASES]

KIND_ARGS = {  # kind -> (op, flags, needs a condition)
    "range": ("range", 0, False),
    "equal": ("equal", 0, False),
    "range_nosimd": ("range", N.BUFF_SCALAR, False),
    "equal_nosimd": ("equal", N.BUFF_SCALAR, False),
    "range_cond": ("range", 0, True),
    "smaller_cond": ("smaller", 0, True),
}


def _preds(x):
    rng = np.random.default_rng(len(x))
    return [float(x.min()) - 1.0, float(np.median(x)), float(x.max()), float(x.max()) * 4 + 1e4] + \
        [float(v) for v in rng.choice(x, 3)] + [float(x[len(x) // 2 & ~1]), float(x[len(x) // 2 | 1])]


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _call_or_panic(fn):
    from oracle import refcpu

    try:
        return fn(), None
    except (N.RefPanic, refcpu.RefPanic) as e:
        return None, e


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_stream_matches_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.slice_encode(x, scale)
    ch = gpu_ctx.encode(x, scale, flags=N.BUFF_SLICE)
    try:
        assert np.array_equal(ch.serialize(), s), name
    finally:
        ch.free()
    ld = gpu_ctx.load(s, scale, flags=N.BUFF_SLICE)
    try:
        assert np.array_equal(ld.serialize(), s), name
        assert _same(ld.materialize(), oracle.slice_decode(s)), name
    finally:
        ld.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_filters_match_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.slice_encode(x, scale)
    ch = gpu_ctx.load(s, scale, flags=N.BUFF_SLICE)
    h = cf.slice_parse(s)
    fl = oracle.encode(x, scale, flip=True) if h["F"] > 0 else None
    fch = gpu_ctx.load(fl, scale, flags=N.BUFF_FLIP) if fl is not None else None
    cond = np.random.default_rng(11).random(len(x)) < 0.4
    cond[:40] = False
    rows = np.flatnonzero(cond)
    cw = cf.to_words(cond)
    try:
        for pred in _preds(x):
            for kind, (op, flags, needs) in KIND_ARGS.items():
                run = (lambda: ch.filter_cond(op, pred, cw, flags=flags)) if needs else \
                    (lambda: ch.filter(op, pred, flags=flags))
                try:
                    words, cnt = oracle.slice_filter(s, scale, kind, pred, rows if needs else None)
                except oracle.RefPanic:
                    with pytest.raises(N.RefPanic):
                        run()
                else:
                    gw, gc = run()
                    assert np.array_equal(gw, words), (name, kind, pred)
                    if not needs:
                        assert gc == cnt, (name, kind, pred)
                # the intended comparison
                sem = cf.slice_filter_bits(s, scale, kind, pred, cond if needs else None, semantic=True)
                gw, _ = (ch.filter_cond(op, pred, cw, flags=flags | N.BUFF_SEMANTIC) if needs else
                         ch.filter(op, pred, flags=flags | N.BUFF_SEMANTIC))
                assert np.array_equal(gw, cf.to_words(sem)), (name, kind, pred, "semantic")
            if fch is None:
                continue
            for kind, op in (("range_with_slice", "range"), ("equal_with_slice", "equal")):
                try:
                    words, cnt = oracle.slice_filter(fl, scale, kind, pred)
                except oracle.RefPanic:
                    with pytest.raises(N.RefPanic):
                        fch.filter(op, pred, flags=N.BUFF_WITH_SLICE)
                    continue
                try:
                    gw, gc = fch.filter(op, pred, flags=N.BUFF_WITH_SLICE)
                except N.Unsupported:
                    # the tail compared after four of more than four planes
                    assert h["F"] // 8 > 4 and h["F"] % 8, (name, kind, pred)
                    continue
                assert np.array_equal(gw, words) and gc == cnt, (name, kind, pred)
    finally:
        ch.free()
        if fch is not None:
            fch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_decodes_sum_max_match_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.slice_encode(x, scale)
    ch = gpu_ctx.load(s, scale, flags=N.BUFF_SLICE)
    h = cf.slice_parse(s)
    try:
        assert _same(ch.materialize(), oracle.slice_decode(s))
        want, err = _call_or_panic(lambda: oracle.slice_decode(s, "bs"))
        if err:
            with pytest.raises(N.RefPanic):
                ch.materialize(flags=N.BUFF_SCALAR)
        else:
            assert _same(ch.materialize(flags=N.BUFF_SCALAR), want)
        with pytest.raises(N.Unsupported):
            ch.materialize(precision=2)
        for frac in (0.3, 0.0):
            cond = np.random.default_rng(4).random(len(x)) < frac
            want, err = _call_or_panic(lambda: oracle.slice_decode(s, "cond", np.flatnonzero(cond)))
            if err:
                with pytest.raises(N.RefPanic):
                    ch.materialize_cond(cf.to_words(cond))
            else:
                assert _same(ch.materialize_cond(cf.to_words(cond)), want)
        exact = float(Fraction(h["n"] * h["base"] + int(h["off"].astype(object).sum()), 1 << h["dlen"])) \
            if h["n"] else 0.0
        vals = cf.slice_decode(s)
        for flags, scalar in ((0, False), (N.BUFF_SCALAR, True)):
            ref, err = _call_or_panic(lambda: oracle.slice_sum(s, scalar))
            if err:
                with pytest.raises(N.RefPanic):
                    ch.sum(flags)
                with pytest.raises(N.RefPanic):
                    ch.max_ex(flags)
                continue
            got, ex, _ = ch.sum(flags)
            assert got == exact and ex == exact
            assert abs(got - ref) <= 1e-12 * max(np.abs(vals).sum(), 1e-300), (name, got, ref)
            mx, cnt, _ = oracle.slice_max(s, scalar)
            gv, _, gc = ch.max_ex(flags)
            assert gv == mx and gc == cnt, (name, scalar)
    finally:
        ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_max_range_matches_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.slice_encode(x, scale)
    ch = gpu_ctx.load(s, scale, flags=N.BUFF_SLICE)
    n = len(x)
    try:
        for a, b, w in ((0, n, n), (0, n - n % 7, 7), (3, 3 + 40 * (max(n - 3, 0) // 40), 40), (0, n, 64),
                        (5, 5, 3), (0, n, 1), (n // 3, n // 3 + 1000 * (2 * n // 3 // 1000), 1000)):
            want, err = _call_or_panic(lambda: oracle.slice_max_range(s, a, b, w))
            if err:
                with pytest.raises(N.RefPanic):
                    ch.max_range(a, b, w)
                continue
            vals, words = ch.max_range(a, b, w)
            assert _same(vals, want[0]) and np.array_equal(words, want[1]), (name, a, b, w)
    finally:
        ch.free()


def test_reference_shaped_entry_points(oracle):
    x = dg.random_walk(50_001, 3, 31)
    bs = buff_amd.BuffSliceCompress(10, 10, 1000)
    s = bs.buff_slice_encode(x)
    assert np.array_equal(s, oracle.slice_encode(x, 1000))
    assert _same(bs.buff_slice_decode(s), oracle.slice_decode(s))
    assert _same(bs.buff_slice_decode_bs(s), oracle.slice_decode(s, "bs"))
    pred = float(np.median(x))
    for fn, kind in (("buff_slice_range_filter", "range"), ("buff_slice_equal_filter", "equal"),
                     ("buff_slice_range_filter_nosimd", "range_nosimd"),
                     ("buff_slice_equal_filter_nosimd", "equal_nosimd")):
        for p in (pred, float(x[123])):
            gw, gc = getattr(bs, fn)(s, p)
            w, c = oracle.slice_filter(s, 1000, kind, p)
            assert np.array_equal(gw, w) and gc == c, fn
    fl = bs.buff_simd256_encode(x)
    for fn, kind in (("buff_simd_range_filter_with_slice", "range_with_slice"),
                     ("buff_simd_equal_filter_with_slice", "equal_with_slice")):
        gw, gc = getattr(bs, fn)(fl, float(x[77]))
        w, c = oracle.slice_filter(fl, 1000, kind, float(x[77]))
        assert np.array_equal(gw, w) and gc == c, fn
    vals = oracle.slice_decode(s)
    for fn, scalar in (("buff_slice_buffsum", False), ("buff_slice_sum", True)):
        assert abs(getattr(bs, fn)(s) - oracle.slice_sum(s, scalar)) <= 1e-12 * np.abs(vals).sum()
    for fn, scalar in (("buff_slice_buffmax", False), ("buff_slice_max", True)):
        mx, cnt, _ = oracle.slice_max(s, scalar)
        assert getattr(bs, fn)(s) == (mx, cnt), fn


def test_constant_column_and_rejections(gpu_ctx, oracle):
    x = np.full(1000, 2.5)
    s = oracle.slice_encode(x, 100)
    ch = gpu_ctx.encode(x, 100, flags=N.BUFF_SLICE)     # F = 0: one plane of 0x80
    try:
        assert np.array_equal(ch.serialize(), s)
        assert _same(ch.materialize(), oracle.slice_decode(s))
        with pytest.raises(N.RefPanic):
            ch.filter("range", 1.0 + 2.5)
        words, cnt = ch.filter("range", 1.0)               # below the base: every row
        assert cnt == 1000
        gv, _, gc = ch.max_ex()
        mx, c, _ = oracle.slice_max(s)
        assert gv == mx and gc == c
    finally:
        ch.free()
    with pytest.raises(N.BuffError):
        gpu_ctx.encode(x, 100, flags=N.BUFF_SLICE | N.BUFF_FLIP)


@pytest.mark.parametrize("F_scale", [(10 ** 4, 4), (10 ** 6, 6)])
def test_at_scale_against_byte_fixed(gpu_ctx, F_scale):
    # 16.8M rows: the slice stream round-trips, its decode is byte_fixed's, and
    # its AVX2 range filter is byte_fixed's over the first 32*floor(n/32) rows
    # when F <= 32
    scale, p = F_scale
    n = 16_777_216 + 333
    x = dg.random_walk(n, p, 17)
    sl = gpu_ctx.encode(x, scale, flags=N.BUFF_SLICE)
    bf = gpu_ctx.encode(x, scale)
    try:
        s = sl.serialize()
        ld = gpu_ctx.load(s, scale, flags=N.BUFF_SLICE)
        try:
            assert np.array_equal(ld.serialize(), s)
        finally:
            ld.free()
        assert _same(sl.materialize(), bf.materialize())
        assert sl.info().stream_bytes == len(s)
        F = sl.info().width
        pred = float(np.median(x))
        gw, _ = sl.filter("range", pred)
        bw, _ = bf.filter("range", pred)
        cover = n // 32
        if F <= 32:
            assert np.array_equal(gw[:cover], bw[:cover])
            assert not gw[cover:].any()
        sw, _ = sl.filter("range", pred, flags=N.BUFF_SEMANTIC)
        assert np.array_equal(sw, bw)
    finally:
        sl.free()
        bf.free()
