"""buff-major (majority / outlier) layout on the GPU vs the oracle, through the C ABI.

Each case of majority_cases.py is encoded on the GPU; the serialized stream must
equal the oracle's byte for byte (snappy bitmaps included), and every query on
the resident chunk (and on a chunk loaded from the oracle's stream) must return
the oracle's answer: bitmaps and counts, decoded doubles bit for bit (or the
reference's panic / empty output), the reference f64 sum bit for bit with its
exact per-column integer sums, the max value, offset and tie count. The
BUFF_SEMANTIC variants are checked against the closed form of the intended
answer (true offsets).
"""
import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg
from tests import closed_form as cf
from tests import majority_cases as mc

pytestmark = pytest.mark.gpu

NAMES = sorted(mc.CASES)


@pytest.fixture(scope="module")
def built(oracle):
    out = {}
    for name in NAMES:
        build, scale = mc.CASES[name]
        x = build()
        out[name] = (x, scale, oracle.major_encode(x, scale))
    return out


def _chunks(gpu_ctx, x, scale, stream):
    enc = gpu_ctx.encode(x, scale, flags=N.BUFF_MAJORITY)
    loaded = gpu_ctx.load(stream, scale, flags=N.BUFF_MAJORITY)
    return enc, loaded


@pytest.mark.parametrize("name", NAMES)
def test_stream_bit_exact(gpu_ctx, built, name):
    x, scale, ref = built[name]
    print("case", name, len(x), flush=True)
    enc, loaded = _chunks(gpu_ctx, x, scale, ref)
    try:
        for ch in (enc, loaded):
            got = ch.serialize()
            assert len(got) == len(ref), name
            assert np.array_equal(got, ref), (name, int(np.argmax(got != ref)))
        h = cf.major_parse(ref)
        inf = enc.info()
        assert inf.sparse_cols == h["S"] and inf.width == h["F"] and inf.base == h["base"]
        for j in range(h["S"]):
            major, nout, last = enc.sparse_col(j)
            assert major == h["majors"][j]
            assert nout == int(h["masks"][j].sum())
            idx = np.flatnonzero(h["masks"][j])
            assert last == (int(idx[-1]) if len(idx) else (1 << 64) - 1)
    finally:
        enc.free()
        loaded.free()


@pytest.mark.parametrize("name", NAMES)
def test_queries_match_oracle(gpu_ctx, oracle, built, name):
    x, scale, ref = built[name]
    print("case", name, len(x), flush=True)
    enc, loaded = _chunks(gpu_ctx, x, scale, ref)
    try:
        for ch in (enc, loaded):
            for pred in mc.predicates(x):
                for op in ("range", "equal"):
                    w_ref, c_ref = oracle.major_filter(ref, scale, op, pred)
                    w, c = ch.filter(op, pred)
                    assert c == c_ref, (name, op, pred)
                    assert np.array_equal(w, w_ref), (name, op, pred)
            s_ref, cs_ref = oracle.major_sum(ref, scale)
            s, exact, cs = ch.sum()
            assert s == s_ref or (np.isnan(s) and np.isnan(s_ref)), name
            assert np.array_equal(cs, cs_ref), name
            v_ref, off_ref, cnt_ref = oracle.major_max(ref, scale)
            assert ch.max_ex() == (v_ref, off_ref, cnt_ref), name
            try:
                d_ref = oracle.major_decode(ref, scale)
            except oracle.RefPanic:
                with pytest.raises(N.RefPanic):
                    ch.materialize()
            else:
                d = ch.materialize()
                assert len(d) == len(d_ref), name
                assert np.array_equal(d.view(np.int64), d_ref.view(np.int64)), name
    finally:
        enc.free()
        loaded.free()


@pytest.mark.parametrize("name", ["config3", "major_zero", "max_quirk_below", "all_sparse", "no_outliers",
                                  "max_in_majority", "max_many_ties"])
def test_semantic_answers(gpu_ctx, built, name):
    """BUFF_SEMANTIC: the intended answers on the true offsets."""
    x, scale, ref = built[name]
    h = cf.major_parse(ref)
    off = np.zeros(h["n"], dtype=np.uint64)
    for k, col in enumerate(h["cols"]):
        off |= col.astype(np.uint64) << np.uint64(h["F"] - 8 * (k + 1))
    off |= h["tail"]
    ch = gpu_ctx.encode(x, scale, flags=N.BUFF_MAJORITY)
    try:
        truth = (np.uint64(h["base"] & ((1 << 64) - 1)) + off).view(np.int64).astype(np.float64) / 2.0 ** h["dlen"]
        d = ch.materialize(flags=N.BUFF_SEMANTIC)
        assert np.array_equal(d.view(np.int64), truth.view(np.int64))
        for pred in mc.predicates(x):
            fp = cf.fixed_pred(pred, scale, h["dlen"])
            for op in ("range", "equal"):
                if fp < h["base"]:
                    want = np.ones(h["n"], bool) if op == "range" else np.zeros(h["n"], bool)
                else:
                    t = fp - h["base"]
                    if t >= 1 << h["F"]:
                        want = np.zeros(h["n"], bool)
                    else:
                        want = off > np.uint64(t) if op == "range" else off == np.uint64(t)
                w, c = ch.filter(op, pred, flags=N.BUFF_SEMANTIC)
                assert np.array_equal(w, cf.to_words(want)), (name, op, pred)
        v, mo, _ = ch.max_ex(flags=N.BUFF_SEMANTIC)
        assert mo == int(off.max())
        _, exact, _ = ch.sum()
        tot = h["n"] * h["base"] + int(off.astype(object).sum())
        assert exact == float(np.float64(tot) / 2.0 ** h["dlen"]) or abs(exact - tot / 2.0 ** h["dlen"]) <= 1e-12 * abs(
            tot / 2.0 ** h["dlen"])
    finally:
        ch.free()


def test_split_double_api(oracle):
    x = mc.case_config3(30000)
    sd = buff_amd.SplitBDDoubleCompress(10, 10, mc.SCALE)
    s = sd.buff_encode_majority(x)
    ref = oracle.major_encode(x, mc.SCALE)
    assert np.array_equal(s, ref)
    assert np.array_equal(sd.buff_decode_majority(s).view(np.int64), oracle.major_decode(ref, mc.SCALE).view(np.int64))
    pred = float(np.quantile(x, 0.9, method="lower"))
    assert sd.buff_range_filter_majority(s, pred)[1] == oracle.major_filter(ref, mc.SCALE, "range", pred)[1]
    assert sd.buff_equal_filter_majority(s, x[17])[1] == oracle.major_filter(ref, mc.SCALE, "equal", x[17])[1]
    assert sd.buff_sum_majority(s) == oracle.major_sum(ref, mc.SCALE)[0]
    assert sd.buff_max_majority(s) == oracle.major_max(ref, mc.SCALE)[0]
    assert len(sd.buff_decode_majority(oracle.major_encode(mc.case_all_sparse(), mc.SCALE))) == 0


def test_reference_panics_and_limits(gpu_ctx, oracle):
    with pytest.raises(N.RefPanic):
        gpu_ctx.encode(mc.case_config3(1999), mc.SCALE, flags=N.BUFF_MAJORITY)
    with pytest.raises(N.RefPanic):
        gpu_ctx.encode(mc.case_narrow_majority(), mc.SCALE, flags=N.BUFF_MAJORITY)
    narrow = mc.case_narrow_plain()
    ch = gpu_ctx.encode(narrow, mc.SCALE, flags=N.BUFF_MAJORITY)
    try:
        assert np.array_equal(ch.serialize(), oracle.major_encode(narrow, mc.SCALE))
    finally:
        ch.free()
    with pytest.raises(N.RefPanic):
        gpu_ctx.load(oracle.major_encode(narrow, mc.SCALE), mc.SCALE, flags=N.BUFF_MAJORITY)
    ch = gpu_ctx.encode(mc.case_config3(5000), mc.SCALE, flags=N.BUFF_MAJORITY)
    try:
        with pytest.raises(N.Unsupported):
            ch.materialize(precision=2)
    finally:
        ch.free()


def test_config3_4M_properties(gpu_ctx, oracle):
    """4,194,304 rows of the config-3 distribution: parity with the oracle on the
    stream, a range filter, the sum and the max."""
    x = dg.outlier_heavy(1 << 22, 3, 4)
    ref = oracle.major_encode(x, mc.SCALE)
    ch = gpu_ctx.encode(x, mc.SCALE, flags=N.BUFF_MAJORITY)
    try:
        assert np.array_equal(ch.serialize(), ref)
        pred = dg.percentile_pred(x, 0.10)
        w, c = ch.filter("range", pred)
        w_ref, c_ref = oracle.major_filter(ref, mc.SCALE, "range", pred)
        assert c == c_ref and np.array_equal(w, w_ref)
        assert ch.sum()[0] == oracle.major_sum(ref, mc.SCALE)[0]
        assert ch.max_ex() == oracle.major_max(ref, mc.SCALE)
    finally:
        ch.free()
