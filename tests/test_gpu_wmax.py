"""Windowed max on the GPU through the C ABI (buff_max_range,
buff_simd.rs:5096-5334): the window maxima and the argmax bitmap against the
oracle bit for bit, with the reference's quirks (full windows only, a panic
when F > 8 and the windows do not tile the range, argmax rows only where its
tail stage runs), and against the closed form with BUFF_SEMANTIC. Windows of
1 row to the whole range, ranges that start and end inside tiles, an empty
range, F < 8, F = 8, F % 8 = 0 and ilen wrapping; 16.8M rows at three window
sizes against the full decode.
"""
import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg
from tests import closed_form as cf

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(41)
    yield "walk4_ragged", dg.random_walk(100_003, 4, 2), 10 ** 4
    yield "uniform2", dg.uniform(65_537, -1000, 1000, 2, 3), 10 ** 2
    yield "outlier3", dg.outlier_heavy(70_001, 3, 4), 10 ** 3
    yield "tail_only", np.round(rng.integers(0, 5, 3001) * 0.1, 1), 10
    yield "F8", np.round(np.linspace(0, 7.9, 4097), 1), 10
    yield "F16_base", np.round(np.linspace(100, 2147.9, 5000), 1), 10
    yield "p6_ilen_wraps", dg.random_walk(20_000, 6, 10) * 1e-3, 10 ** 6
    yield "two_rows", np.array([3.25, 3.5]), 10 ** 2


CASES = list(_cases())
IDS = [c[0] for c in CASES]


def _spans(n, seed):
    rng = np.random.default_rng(seed)
    out = [(0, n, n), (0, n, 1), (0, n - n % 7, 7), (0, 0, 5), (n // 2, n // 2 + 1, 1)]
    for w in (3, 64, 100, 1000, 1024, 5000):
        a = int(rng.integers(0, max(n // 3, 1)))
        b = int(rng.integers(a, n + 1))
        out.append((a, b, w))                          # usually a partial last window
        out.append((a, a + (b - a) // w * w, w))       # exact tiling
    return out


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_max_range_matches_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.encode(x, scale)
    ch = gpu_ctx.encode(x, scale)
    h = cf.parse(s)
    try:
        for a, b, w in _spans(len(x), len(x)):
            try:
                offs, vals, words = oracle.max_range(s, a, b, w)
            except oracle.RefPanic:
                with pytest.raises(N.RefPanic):
                    ch.max_range(a, b, w)
            else:
                gv, gw = ch.max_range(a, b, w)
                assert np.array_equal(gv.view(np.uint64), vals.view(np.uint64)), (name, a, b, w)
                assert np.array_equal(gw, words), (name, a, b, w)
            sm, sarg = cf.max_range(s, a, b, w, semantic=True)
            gv, gw = ch.max_range(a, b, w, flags=N.BUFF_SEMANTIC)
            want = (sm.astype(np.int64) + h["base"]).astype(np.float64) / float(1 << h["dlen"])
            assert np.array_equal(gv.view(np.uint64), want.view(np.uint64)), (name, a, b, w, "semantic")
            assert np.array_equal(gw, cf.to_words(sarg)), (name, a, b, w, "semantic")
        with pytest.raises(N.BuffError):
            ch.max_range(0, len(x), 0)
        with pytest.raises(N.BuffError):
            ch.max_range(0, len(x) + 1, 4)
    finally:
        ch.free()


@pytest.mark.parametrize("window", [7, 1000, 1 << 20])
def test_max_range_at_scale(gpu_ctx, window):
    n = 16_777_216 + 333
    x = dg.random_walk(n, 4, 9)
    ch = gpu_ctx.encode(x, 10 ** 4)
    try:
        full = ch.materialize()
        a, b = 1234, n - 77
        vals, words = ch.max_range(a, b, window, flags=N.BUFF_SEMANTIC)
        seg = full[a:b]
        starts = np.arange(0, len(seg), window)
        want = np.maximum.reduceat(seg, starts)
        assert np.array_equal(vals.view(np.uint64), want.view(np.uint64))
        arg = np.zeros(n, bool)
        arg[a:b] = seg == np.repeat(want, np.diff(np.append(starts, len(seg))))
        assert np.array_equal(words, cf.to_words(arg))
    finally:
        ch.free()


@pytest.mark.parametrize("case", ["config3", "no_majority", "all_sparse", "no_outliers"])
def test_majority_form_matches_oracle(gpu_ctx, oracle, case):
    # buff_max_majority_range: no window and no argmax row when the chunk has
    # sparse columns, the byte_fixed result when it has none
    from tests import majority_cases as mc

    build, scale = mc.CASES[case]
    x = build()
    s = oracle.major_encode(x, scale)
    ch = gpu_ctx.encode(x, scale, flags=N.BUFF_MAJORITY)
    try:
        sparse = ch.info().sparse_cols > 0
        for a, b, w in _spans(len(x), 3 + len(x)):
            try:
                offs, vals, words = oracle.max_range(s, a, b, w, major=True)
            except oracle.RefPanic:
                with pytest.raises(N.RefPanic):
                    ch.max_range(a, b, w)
                continue
            gv, gw = ch.max_range(a, b, w)
            assert np.array_equal(gv.view(np.uint64), vals.view(np.uint64)), (case, a, b, w)
            assert np.array_equal(gw, words), (case, a, b, w)
            if sparse:
                assert len(gv) == 0 and not gw.any()
                with pytest.raises(N.Unsupported):
                    ch.max_range(a, b, w, flags=N.BUFF_SEMANTIC)
    finally:
        ch.free()
