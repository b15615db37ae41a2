"""Filter -> project on the GPU through the C ABI: the conditional range filter
(buff_range_filter_condition, buff_simd.rs:4683-4921) and the conditional
decode (buff_decode_condition, buff_simd.rs:1528-1712) against the oracle, bit
for bit, with the reference's quirks (no rows for a predicate below the base,
the low-byte target when F < 8, the tail bits shifted by r in the decode, the
panic on an empty condition), and against the closed forms with
BUFF_SEMANTIC. Conditions are random row sets of every density, the last row
alone, every row, no row, and an earlier filter's result chained on the device.
"""
import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg
from tests import closed_form as cf

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(31)
    yield "walk4_ragged", dg.random_walk(100_003, 4, 2), 10 ** 4
    yield "uniform2", dg.uniform(65_537, -1000, 1000, 2, 3), 10 ** 2
    yield "outlier3", dg.outlier_heavy(70_001, 3, 4), 10 ** 3
    yield "tail_only", np.round(rng.integers(0, 5, 3001) * 0.1, 1), 10
    yield "F16_base", np.round(np.linspace(100, 2147.9, 5000), 1), 10
    yield "F8", np.round(np.linspace(0, 7.9, 4097), 1), 10
    yield "p6_ilen_wraps", dg.random_walk(20_000, 6, 10) * 1e-3, 10 ** 6
    yield "two_rows", np.array([3.25, 3.5]), 10 ** 2
    yield "n1025", dg.random_walk(1025, 2, 12), 10 ** 2


CASES = list(_cases())
IDS = [c[0] for c in CASES]


def _conds(n, seed):
    rng = np.random.default_rng(seed)
    for frac in (0.001, 0.05, 0.5, 0.97):
        yield "rand%g" % frac, rng.random(n) < frac
    last = np.zeros(n, bool)
    last[-1] = True
    yield "last_row", last
    yield "all", np.ones(n, bool)
    yield "none", np.zeros(n, bool)
    block = np.zeros(n, bool)
    block[n // 3: n // 3 + 2048] = True   # whole tiles with and without rows
    yield "block", block


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_filter_cond_matches_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.encode(x, scale)
    ch = gpu_ctx.encode(x, scale)
    try:
        rng = np.random.default_rng(len(x))
        preds = list(rng.choice(x, 3)) + [float(x.min()) - 1.0, float(np.median(x)), float(x.max()),
                                          float(abs(x).max()) * 8 + 1e5]
        for cname, cond in _conds(len(x), 7):
            cw = cf.to_words(cond)
            rows = np.flatnonzero(cond)
            for pred in preds:
                words, cnt = ch.filter_cond("range", pred, cw)
                ref, rcnt = oracle.filter_cond(s, scale, pred, rows)
                assert np.array_equal(words, ref), (name, cname, pred)
                assert cnt == rcnt
                sem, _ = ch.filter_cond("range", pred, cw, flags=N.BUFF_SEMANTIC)
                want = cf.filter_cond_bits(s, scale, "range", pred, cond, semantic=True)
                assert np.array_equal(sem, cf.to_words(want)), (name, cname, pred, "semantic")
                for flags in (0, N.BUFF_SEMANTIC):
                    eq, _ = ch.filter_cond("equal", pred, cw, flags=flags)
                    want = cf.filter_bits(s, scale, "equal", pred, semantic=bool(flags)) & cond
                    assert np.array_equal(eq, cf.to_words(want)), (name, cname, pred, "equal", flags)
    finally:
        ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_materialize_cond_matches_oracle(gpu_ctx, oracle, name, x, scale):
    s = oracle.encode(x, scale)
    ch = gpu_ctx.encode(x, scale)
    try:
        for cname, cond in _conds(len(x), 9):
            cw = cf.to_words(cond)
            rows = np.flatnonzero(cond)
            if len(rows) == 0:
                with pytest.raises(N.RefPanic):
                    ch.materialize_cond(cw)
                assert len(ch.materialize_cond(cw, flags=N.BUFF_SEMANTIC)) == 0
                continue
            got = ch.materialize_cond(cw)
            ref = oracle.decode_cond(s, scale, rows)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (name, cname)
            sem = ch.materialize_cond(cw, flags=N.BUFF_SEMANTIC)
            assert np.array_equal(sem.view(np.uint64), cf.decode_cond(s, cond, semantic=True).view(np.uint64))
    finally:
        ch.free()


def test_chained_on_device(gpu_ctx, oracle):
    # filter -> conditional filter -> conditional decode, the bitmaps never leaving the GPU
    x = dg.random_walk(300_001, 4, 5)
    scale = 10 ** 4
    s = oracle.encode(x, scale)
    n = len(x)
    lo, hi = float(np.quantile(x, 0.3, method="lower")), float(np.quantile(x, 0.6, method="lower"))
    ch = gpu_ctx.encode(x, scale)
    nw = (n + 63) // 64
    a, b, out = gpu_ctx.alloc(nw * 8), gpu_ctx.alloc(nw * 8), gpu_ctx.alloc(n * 8)
    try:
        ch.filter_device("range", lo, a.ptr)
        ch.filter_cond_device("range", hi, a.ptr, b.ptr, flags=N.BUFF_SEMANTIC)   # lo < x and hi < x
        first = np.unpackbits(a.download(np.uint64, nw).view(np.uint8), bitorder="little")[:n].astype(bool)
        both = np.unpackbits(b.download(np.uint64, nw).view(np.uint8), bitorder="little")[:n].astype(bool)
        assert np.array_equal(both, cf.filter_bits(s, scale, "range", hi) & first)
        rows = ch.materialize_cond_device(b.ptr, out.ptr, n, flags=N.BUFF_SEMANTIC)
        assert rows == int(both.sum())
        got = out.download(np.float64, rows)
        assert np.array_equal(got.view(np.uint64), cf.decode_cond(s, both, semantic=True).view(np.uint64))
        # a buffer too small: the row count comes back with the error, nothing past cap is written
        marker = np.full(n, -7.0)
        out.upload(marker)
        with pytest.raises(N.BuffError):
            ch.materialize_cond_device(b.ptr, out.ptr, rows // 2, flags=N.BUFF_SEMANTIC)
        tail = out.download(np.float64, n)[rows // 2:]
        assert np.all(tail == -7.0)
    finally:
        for buf in (a, b, out):
            buf.free()
        ch.free()


def test_all_rows_is_the_plain_path_at_scale(gpu_ctx):
    # a condition of every row: the conditional filter equals the filter and the
    # conditional decode (semantic) equals the full decode, bit for bit
    n = 16_777_216 + 333
    x = dg.random_walk(n, 4, 8)
    ch = gpu_ctx.encode(x, 10 ** 4)
    nw = (n + 63) // 64
    ones, w1, w2, v1, v2 = (gpu_ctx.alloc(nw * 8), gpu_ctx.alloc(nw * 8), gpu_ctx.alloc(nw * 8),
                            gpu_ctx.alloc(n * 8), gpu_ctx.alloc(n * 8))
    try:
        full = np.full(nw, np.uint64(0xFFFFFFFFFFFFFFFF))
        full[-1] = np.uint64((1 << (n % 64)) - 1) if n % 64 else full[-1]
        ones.upload(full)
        pred = float(np.median(x))
        ch.filter_device("range", pred, w1.ptr)
        ch.filter_cond_device("range", pred, ones.ptr, w2.ptr)
        assert np.array_equal(w1.download(np.uint64, nw), w2.download(np.uint64, nw))
        ch.materialize_device(v1.ptr)
        rows = ch.materialize_cond_device(ones.ptr, v2.ptr, n, flags=N.BUFF_SEMANTIC)
        assert rows == n
        assert np.array_equal(v1.download(np.float64, n).view(np.uint64), v2.download(np.float64, n).view(np.uint64))
    finally:
        for buf in (ones, w1, w2, v1, v2):
            buf.free()
        ch.free()


@pytest.mark.parametrize("case", ["config3", "no_majority", "all_sparse"])
def test_majority_forms_match_oracle(gpu_ctx, oracle, case):
    # buff_range_filter_majority_condition / buff_major_decode_condition: nothing
    # when the chunk has sparse columns, the byte_fixed forms (tail in place, no
    # panic on an empty condition) when it has none
    from tests import majority_cases as mc

    build, scale = mc.CASES[case]
    x = build()
    s = oracle.major_encode(x, scale)
    ch = gpu_ctx.encode(x, scale, flags=buff_amd.BUFF_MAJORITY)
    try:
        sparse = ch.info().sparse_cols > 0
        for cname, cond in _conds(len(x), 13):
            cw = cf.to_words(cond)
            rows = np.flatnonzero(cond)
            for pred in mc.predicates(x)[:6]:
                words, cnt = ch.filter_cond("range", pred, cw)
                ref, rcnt = oracle.filter_cond(s, scale, pred, rows, major=True)
                assert np.array_equal(words, ref) and cnt == rcnt, (case, cname, pred)
            got = ch.materialize_cond(cw)
            ref = oracle.decode_cond(s, scale, rows, major=True)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (case, cname)
            # the intended values: the semantic full decode on the condition rows
            full_vals = ch.materialize(flags=N.BUFF_SEMANTIC)
            sem = ch.materialize_cond(cw, flags=N.BUFF_SEMANTIC)
            assert np.array_equal(sem.view(np.uint64), full_vals[cond].view(np.uint64)), (case, cname, sparse)
            # the intended answer and the equality form: the chunk's own filter on the condition rows
            for pred in mc.predicates(x)[:4]:
                for op, flags in (("range", N.BUFF_SEMANTIC), ("equal", 0), ("equal", N.BUFF_SEMANTIC)):
                    full, _ = ch.filter(op, pred, flags=flags)
                    got, cnt = ch.filter_cond(op, pred, cw, flags=flags)
                    assert np.array_equal(got, full & cw), (case, cname, pred, op, flags)
                    assert cnt == int(np.unpackbits(got.view(np.uint8)).sum())
    finally:
        ch.free()
