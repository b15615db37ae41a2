"""Pins the oracle's byte-slice restatement (oracle/refslice.inc: buff_slice.rs:
222-1970 and the buff-simd `_with_slice` filters, buff_simd.rs:3736-4025)
against the closed forms in closed_form.py, which are written from the stream
format and the comparison arithmetic rather than the reference's loops, and
against the byte_fixed oracle (same header, same offsets). Cases cover F = 0,
F < 8, F = 8, F % 8 = 0, F > 32 (more planes than the filters read), n < 32,
n % 32 != 0, and predicates below the base, above the maximum and on stored
values.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import refcpu as R
from tests import closed_form as cf


def _cases():
    rng = np.random.default_rng(5)
    yield "walk2_ragged", np.round(np.cumsum(rng.normal(0, 1, 3001)), 2), 100
    yield "walk4", np.round(np.cumsum(rng.normal(0, 1, 4096)), 4), 10 ** 4
    yield "F33_p6", np.round(np.cumsum(rng.normal(0, 40, 2000)), 6), 10 ** 6     # 5 planes
    yield "F26_p6", np.round(np.cumsum(rng.normal(0, 1, 2048)), 6), 10 ** 6      # 4 planes
    yield "F44_p9", np.round(np.cumsum(rng.normal(0, 200, 1500)), 9), 10 ** 9    # 6 planes and a tail
    # F = 33: pairs of offsets (2k, 2k + 1) that agree on the four planes the
    # filters read and differ in the fifth
    v = np.arange(0, 3000, 1.5)
    yield "F33_pairs", np.stack([v, v + 2.0 ** -21], axis=1).reshape(-1)[:4000], 10 ** 6
    # F = 45 (5 planes + 5 tail bits): offsets t + (m << 5) that differ only in
    # the fifth plane, which the with_slice filters skip before the tail
    k = np.arange(0, 16000, 7.0)
    yield "F45_mid", np.concatenate([k, k + 3 * 2.0 ** -26, k[::3] + 3 * 2.0 ** -26 + 2.0 ** -31]), 10 ** 9
    yield "F_lt8", np.round(rng.integers(0, 9, 500) * 0.1, 1), 10
    yield "F8", np.round(np.linspace(0, 7.9, 777), 1), 10
    yield "F16", np.round(np.linspace(100, 2147.9, 1500), 1), 10
    yield "const", np.full(100, 3.25), 100
    yield "n31", np.round(rng.normal(0, 10, 31), 2), 100
    yield "n32", np.round(rng.normal(0, 10, 32), 2), 100
    yield "p3_ilen_wraps", np.round(rng.random(700) * 0.004, 3), 1000


CASES = list(_cases())
IDS = [c[0] for c in CASES]


def _preds(x):
    rng = np.random.default_rng(len(x))
    return [float(x.min()) - 1.0, float(np.median(x)), float(x.max()), float(x.max()) * 4 + 1e4] + \
        [float(v) for v in rng.choice(x, 3)] + [float(x[len(x) // 2 & ~1]), float(x[len(x) // 2 | 1])]


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_encode_and_decode(name, x, scale):
    s = R.slice_encode(x, scale)
    assert np.array_equal(s, cf.slice_encode(x, scale)), name
    h = cf.slice_parse(s)
    if h["F"] > 0:
        # the byte_fixed oracle stores the same header and offsets
        b = R.encode(x, scale)
        assert np.array_equal(b[:20], s[:20])
        assert np.array_equal(cf.parse(b)["off"], h["off"])
    assert _same(R.slice_decode(s), cf.slice_decode(s))
    try:
        want = cf.slice_decode(s, "bs")
    except cf.Panic:
        with pytest.raises(R.RefPanic):
            R.slice_decode(s, "bs")
    else:
        assert _same(R.slice_decode(s, "bs"), want)
    rows = np.flatnonzero(np.random.default_rng(3).random(len(x)) < 0.3)
    for rr in (rows, rows[:0], np.array([len(x) - 1])):
        try:
            want = cf.slice_decode(s, "cond", rr)
        except cf.Panic:
            with pytest.raises(R.RefPanic):
                R.slice_decode(s, "cond", rr)
        else:
            assert _same(R.slice_decode(s, "cond", rr), want)


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_sum_and_max(name, x, scale):
    s = R.slice_encode(x, scale)
    h = cf.slice_parse(s)
    exact = float(Fraction(h["n"] * h["base"] + int(h["off"].astype(object).sum()), 1 << h["dlen"])) \
        if h["n"] else 0.0
    vals = cf.slice_decode(s)
    serial = 0.0
    for v in vals:
        serial += v
    assert R.slice_sum(s) == serial                    # the running f64 sum, in row order
    assert abs(serial - exact) <= 1e-12 * max(np.abs(vals).sum(), 1e-300)
    for scalar in (False, True):
        try:
            mx, arg = cf.slice_max(s, scalar)
        except cf.Panic:
            with pytest.raises(R.RefPanic):
                R.slice_max(s, scalar)
            with pytest.raises(R.RefPanic):
                R.slice_sum(s, scalar)
            continue
        if scalar:
            assert R.slice_sum(s, True) == serial
        got, cnt, words = R.slice_max(s, scalar)
        assert got == mx and cnt == int(arg.sum())
        assert np.array_equal(words, cf.to_words(arg))


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_max_range(name, x, scale):
    s = R.slice_encode(x, scale)
    n = len(x)
    for a, b, w in ((0, n, n), (0, n - n % 7, 7), (3, 3 + 40 * (max(n - 3, 0) // 40), 40), (0, n, 64), (5, 5, 3),
                    (0, n, 1)):
        try:
            wm, arg = cf.slice_max_range(s, a, b, w)
        except cf.Panic:
            with pytest.raises(R.RefPanic):
                R.slice_max_range(s, a, b, w)
            continue
        vals, words = R.slice_max_range(s, a, b, w)
        assert _same(vals, wm), (name, a, b, w)
        assert np.array_equal(words, cf.to_words(arg)), (name, a, b, w)


def test_remove_range_at_word_starts():
    # windows starting on a 32-row boundary take remove_range(s, s), whose
    # release-mode shift clears that whole block: rows of the window not yet
    # read, so the result is unchanged (a rising run makes every row a new max)
    x = np.round(np.arange(256) * 0.01, 2)
    s = R.slice_encode(x, 100)
    for w in (32, 64, 1, 5):
        vals, words = R.slice_max_range(s, 0, 256 - 256 % w, w)
        wm, arg = cf.slice_max_range(s, 0, 256 - 256 % w, w)
        assert _same(vals, wm) and np.array_equal(words, cf.to_words(arg)), w


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_filters(name, x, scale):
    s = R.slice_encode(x, scale)
    fl = R.encode(x, scale, flip=True) if cf.slice_parse(s)["F"] > 0 else None
    cond = np.random.default_rng(11).random(len(x)) < 0.4
    cond[:40] = False                 # a whole condition word without rows
    rows = np.flatnonzero(cond)
    for kind in R.SLICE_KINDS:
        stream = s
        if kind.endswith("with_slice"):
            if fl is None:
                continue
            stream = fl
        c = cond if kind.endswith("_cond") else None
        for pred in _preds(x):
            try:
                want = cf.slice_filter_bits(stream, scale, kind, pred, c)
            except cf.Panic:
                with pytest.raises(R.RefPanic):
                    R.slice_filter(stream, scale, kind, pred, rows if c is not None else None)
                continue
            words, cnt = R.slice_filter(stream, scale, kind, pred, rows if c is not None else None)
            assert np.array_equal(words, cf.to_words(want)), (name, kind, pred)
            assert cnt == int(want.sum()), (name, kind, pred)


def test_filters_agree_with_numpy_on_a_plain_column():
    # the closed forms themselves, on a column where every variant is the plain
    # comparison: F <= 32, n % 32 == 0, predicate inside the column
    x = np.round(np.random.default_rng(2).normal(0, 30, 2048), 2)
    s = R.slice_encode(x, 100)
    fl = R.encode(x, 100, flip=True)
    v = cf.slice_decode(s)
    for pred in (float(np.median(x)), float(x[17])):
        p = cf.fixed_pred(pred, 100, 8) / 2.0 ** 8
        assert np.array_equal(cf.slice_filter_bits(s, 100, "range", pred), v > p)
        assert np.array_equal(cf.slice_filter_bits(s, 100, "equal_nosimd", pred), v == p)
        assert np.array_equal(cf.slice_filter_bits(fl, 100, "range_with_slice", pred), v > p)
        cond = v > 0
        assert np.array_equal(cf.slice_filter_bits(s, 100, "smaller_cond", pred, cond), (v < p) & cond)
