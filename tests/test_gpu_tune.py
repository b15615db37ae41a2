"""Launch-shape tuning never changes an answer.

A chunk is tuned (every candidate shape of the range filter, sum, max and full
materialization is run on it), and every query on it, and on a second chunk of
the same size class and byte planes that now uses the tuned shapes, must still
equal the oracle bit for bit.
"""
import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg

pytestmark = pytest.mark.gpu


def _check(ch, oracle, x, scale):
    ref = oracle.encode(x, scale)
    assert np.array_equal(ch.serialize(), ref)
    assert np.array_equal(ch.materialize().view(np.uint64), oracle.decode(ref, scale).view(np.uint64))
    for pred in (float(np.quantile(x, 0.9, method="lower")), float(np.median(x)), float(x[17]), float(x.max())):
        for op in ("range", "equal"):
            w_ref, c_ref, _ = oracle.buff_filter(ref, scale, op, pred)
            w, c = ch.filter(op, pred)
            assert c == c_ref and np.array_equal(w, w_ref), (op, pred)
    s_ref, _ = oracle.byte_fixed_sum(ref, scale)
    assert ch.sum()[0] == s_ref
    m_ref, off_ref, _ = oracle.byte_fixed_max(ref, scale)
    assert ch.max() == (m_ref, off_ref)


def test_tuned_shapes_keep_results(gpu_ctx, oracle):
    x = dg.random_walk(700_001, 4, 31)
    ch = gpu_ctx.encode(x, 10 ** 4)
    try:
        choices = ch.tune(float(np.quantile(x, 0.9, method="lower")))
        assert set(choices) == {"range_filter", "sum", "max", "materialize"}
        for c in choices.values():
            assert c["waves_per_wg"] in (2, 4) and c["max_wg"] in (0, 1024, 2048, 4096, 8192, 16384)
            assert 0 < c["best_us"] <= c["default_us"]
        _check(ch, oracle, x, 10 ** 4)
    finally:
        ch.free()
    y = dg.random_walk(650_000, 4, 32)   # same size class and byte planes: runs with the tuned shapes
    ch2 = gpu_ctx.encode(y, 10 ** 4)
    try:
        _check(ch2, oracle, y, 10 ** 4)
    finally:
        ch2.free()


def test_tune_rejects_unsupported(gpu_ctx):
    x = dg.outlier_heavy(50_000, 3, 5)
    ch = gpu_ctx.encode(x, 10 ** 3, flags=N.BUFF_MAJORITY)
    try:
        with pytest.raises(N.Unsupported):
            ch.tune(100.0)
    finally:
        ch.free()
