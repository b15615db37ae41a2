"""HIP path vs the oracle, bit for bit, through the C ABI (libbuffgpu.so).

Cases cover the BASELINE workloads at sizes the oracle finishes in seconds,
plus the edge cases of the reference: ragged lengths (not multiples of 32, 64,
1024), a single row, the empty column, F < 8 (tail only), F % 8 == 0 (no tail:
equality answer left in rb1, sum without the base term), row 0 holding the
unique maximum (`else if` max), predicates below the base and above the max,
and the flipped buff-simd layout.
"""
import numpy as np
import pytest

from buff_amd import datagen as dg
from tests import closed_form as cf

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(21)
    yield "cfg1_walk_1M", dg.random_walk(1 << 20, 4, 1), 10 ** 4
    yield "walk4_ragged", dg.random_walk(100_003, 4, 2), 10 ** 4
    yield "uniform2", dg.uniform(65_537, -1000, 1000, 2, 3), 10 ** 2
    yield "outlier3", dg.outlier_heavy(200_001, 3, 4), 10 ** 3
    yield "tail_only", np.round(rng.integers(0, 5, 3001) * 0.1, 1), 10
    yield "F16_base", np.round(np.linspace(100, 2147.9, 5000), 1), 10
    yield "F8", np.round(np.linspace(0, 7.9, 4097), 1), 10
    yield "F8_base", np.round(np.linspace(100, 107.9, 4097), 1), 10
    yield "walk4_r6", dg.random_walk(200_001, 4, 15) * 64, 10 ** 4   # F = 30: q < r reads at p = 3
    yield "first_is_max", np.concatenate([[50.0], np.round(np.linspace(0, 10, 70_000), 2)]), 10 ** 2
    yield "negative", dg.random_walk(33_333, 3, 9) - 500, 10 ** 3
    yield "p6_ilen_wraps", dg.random_walk(20_000, 6, 10) * 1e-3, 10 ** 6
    yield "two_rows", np.array([3.25, 3.5]), 10 ** 2
    yield "n63", dg.random_walk(63, 4, 11), 10 ** 4
    yield "n1025", dg.random_walk(1025, 2, 12), 10 ** 2
    yield "p5", dg.random_walk(50_000, 5, 13), 10 ** 5
    yield "p1", dg.random_walk(50_000, 1, 14), 10


CASES = list(_cases())
IDS = [c[0] for c in CASES]


def _preds(x):
    rng = np.random.default_rng(len(x))
    p = list(rng.choice(x, 4))
    p += [float(np.quantile(x, 0.9, method="lower")), float(np.quantile(x, 0.999, method="lower")),
          float(np.median(x)), float(x.min()) - 1.0, float(x.min()), float(x.max()), float(x.max()) + 0.5,
          float(abs(x).max()) * 8 + 1e5]
    return p


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_encode_stream_bit_exact(gpu_ctx, oracle, name, x, scale):
    for flip in (False, True):
        ref = oracle.encode(x, scale, flip=flip)
        ch = gpu_ctx.encode(x, scale, flags=1 if flip else 0)
        got = ch.serialize()
        ch.free()
        assert len(got) == len(ref), name
        assert np.array_equal(got, ref), (name, flip, int(np.argmax(got != ref)))


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_load_serialize_roundtrip(gpu_ctx, oracle, name, x, scale):
    for flip in (False, True):
        ref = oracle.encode(x, scale, flip=flip)
        ch = gpu_ctx.load(ref, scale, flags=1 if flip else 0)
        assert np.array_equal(ch.serialize(), ref), name
        ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_queries_bit_exact(gpu_ctx, oracle, name, x, scale):
    ref = oracle.encode(x, scale)
    ch = gpu_ctx.load(ref, scale)
    # materialize (byte_fixed_decode)
    assert np.array_equal(ch.materialize().view(np.uint64), oracle.decode(ref, scale).view(np.uint64)), name
    # filters
    for pred in _preds(x):
        for op in ("range", "equal"):
            w_ref, c_ref, _ = oracle.buff_filter(ref, scale, op, pred)
            w, c = ch.filter(op, pred)
            assert c == c_ref, (name, op, pred)
            assert np.array_equal(w, w_ref), (name, op, pred)
    # sum: reference fold bit for bit, exact per-sub-column sums
    s_ref, ps_ref = oracle.byte_fixed_sum(ref, scale)
    s, s_exact, ps = ch.sum()
    assert s == s_ref or (np.isnan(s) and np.isnan(s_ref)), (name, s, s_ref)
    h = cf.parse(ref)
    k = h["P"] + (1 if h["r"] else 0)
    assert list(ps[:k]) == list(ps_ref[:k])
    tot, _ = cf.exact_sum(ref)
    assert abs(s_exact - float(tot)) <= 1e-12 * max(1.0, abs(float(tot)))
    # max
    m_ref, off_ref, _ = oracle.byte_fixed_max(ref, scale)
    m, off = ch.max()
    assert off == off_ref and m == m_ref, (name, m, m_ref)
    ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_reduced_precision(gpu_ctx, oracle, name, x, scale):
    import buff_amd

    ref = oracle.encode(x, scale)
    ch = gpu_ctx.load(ref, scale)
    for p in range(0, 7):
        try:
            want = oracle.decode_with_precision(ref, scale, p)
            err = None
        except oracle.RefPanic:
            want, err = None, buff_amd.RefPanic
        except oracle.RefUnsupported:
            want, err = None, buff_amd.Unsupported
        if err is None:
            h = cf.parse(ref)
            R = cf.precision_R(h, p)
            if R // 8 > h["P"]:
                err = buff_amd.Unsupported
        if err is not None:
            with pytest.raises(err):
                ch.materialize(p)
        else:
            got = ch.materialize(p)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, p)
        # semantic: top R bits, re-scaled
        try:
            sem_want = cf.decode_semantic_precision(ref, p)
        except (ValueError, KeyError):
            sem_want = cf.decode(ref)
        got = ch.materialize(p, flags=buff_amd.BUFF_SEMANTIC)
        assert np.array_equal(got.view(np.uint64), sem_want.view(np.uint64)), (name, "semantic", p)
    ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_sum_with_precision(gpu_ctx, oracle, name, x, scale):
    # byte_residue_sum_with_precision: the reference fold bit for bit (or the
    # same panic / unsupported branch), the u64 sums it reads, and the exact
    # sum of the truncated values against the closed form
    import buff_amd
    from buff_amd.split_double import SplitBDDoubleCompress

    ref = oracle.encode(x, scale)
    h = cf.parse(ref)
    ch = gpu_ctx.load(ref, scale)
    sd = SplitBDDoubleCompress(10, 10, scale)
    for p in range(0, 7):
        try:
            want, ps_want = oracle.sum_with_precision(ref, scale, p)
            err = None
        except oracle.RefPanic:
            err = buff_amd.RefPanic
        except oracle.RefUnsupported:
            err = buff_amd.Unsupported
        if err is None and cf.precision_R(h, p) // 8 > h["P"]:
            err = buff_amd.Unsupported   # whole bytes read past the planes of a tiny column
        if err is not None:
            with pytest.raises(err):
                ch.sum_precision(p)
            # the reference-shaped entry point parses its input as a residue
            # stream: a byte_fixed stream with ilen > 15 has no outlier section,
            # its first plane byte is taken for the column count and the parse fails
            with pytest.raises(err if h["ilen"] <= 15 else buff_amd.BuffError):
                sd.byte_residue_sum_with_precision(ref, p)
        else:
            got, exact, ps = ch.sum_precision(p)
            assert got == want or (np.isnan(got) and np.isnan(want)), (name, p, got, want)
            R = cf.precision_R(h, p)
            k = R // 8 + (1 if R % 8 else 0)
            assert list(ps[:k]) == list(ps_want[:k]), (name, p)
            sem = cf.sum_semantic_precision(ref, p)
            assert abs(exact - float(sem)) <= 1e-12 * max(1.0, abs(float(sem))), (name, p)
            assert sd.byte_residue_sum_with_precision(ref, p) == got
        # BUFF_SEMANTIC: the exact truncated sum (every bit when the reference would panic)
        try:
            sem = cf.sum_semantic_precision(ref, p)
        except (ValueError, KeyError):
            sem = cf.exact_sum(ref)[0]
        r2, e2, _ = ch.sum_precision(p, flags=buff_amd.BUFF_SEMANTIC)
        assert r2 == e2 and abs(e2 - float(sem)) <= 1e-12 * max(1.0, abs(float(sem))), (name, "semantic", p)
    ch.free()


@pytest.mark.parametrize("name,x,scale", CASES, ids=IDS)
def test_semantic_filters(gpu_ctx, oracle, name, x, scale):
    import buff_amd

    ref = oracle.encode(x, scale)
    ch = gpu_ctx.load(ref, scale)
    for pred in _preds(x):
        for op in ("range", "equal"):
            w, c = ch.filter(op, pred, flags=buff_amd.BUFF_SEMANTIC)
            m = cf.filter_bits(ref, scale, op, pred, semantic=True)
            assert np.array_equal(w, cf.to_words(m)), (name, op, pred)
            assert c == int(m.sum())
    ch.free()


def test_simd_driver_entry_points(oracle):
    """buff_sd_* (reference method shapes) against the oracle."""
    import buff_amd

    x = dg.random_walk(100_000, 4, 31)
    sd = buff_amd.SplitBDDoubleCompress(10, 10, 10 ** 4)
    s = sd.byte_fixed_encode(x)
    assert np.array_equal(s, oracle.encode(x, 10 ** 4))
    sf = sd.buff_simd256_encode(x)
    assert np.array_equal(sf, oracle.encode(x, 10 ** 4, flip=True))
    assert np.array_equal(sd.byte_fixed_decode(s).view(np.uint64), oracle.decode(s, 10 ** 4).view(np.uint64))
    assert np.array_equal(sd.buff_simd256_decode(sf).view(np.uint64), oracle.decode(s, 10 ** 4).view(np.uint64))
    pred = 1.1509
    for op, fn in (("range", sd.buff_range_filter), ("equal", sd.buff_equal_filter)):
        w_ref, c_ref, _ = oracle.buff_filter(s, 10 ** 4, op, pred)
        w, c = fn(s, pred)
        assert c == c_ref and np.array_equal(w, w_ref)
    for op, fn in (("range", sd.buff_simd_range_filter), ("equal", sd.buff_simd_equal_filter)):
        w_ref, c_ref, _ = oracle.buff_filter(s, 10 ** 4, op, pred)
        w, c = fn(sf, pred)
        assert c == c_ref and np.array_equal(w, w_ref)
    assert sd.byte_fixed_sum(s) == oracle.byte_fixed_sum(s, 10 ** 4)[0]
    assert sd.byte_fixed_max(s) == oracle.byte_fixed_max(s, 10 ** 4)[0]
    got = sd.byte_residue_decode_with_precision(s, 2)
    assert np.array_equal(got.view(np.uint64), oracle.decode_with_precision(s, 10 ** 4, 2).view(np.uint64))


def test_errors_are_status_codes(gpu_ctx):
    import buff_amd

    with pytest.raises(buff_amd.RefPanic):
        gpu_ctx.encode(np.full(100, 2.5), 10 ** 2)           # F = 0
    with pytest.raises(buff_amd.RefPanic):
        gpu_ctx.encode(np.array([1.0, 2.0]), 1)               # precision 0 not in PRECISION_MAP
    bad = np.zeros(10, dtype=np.uint8)
    with pytest.raises(buff_amd.BuffError):
        gpu_ctx.load(bad, 10 ** 4)                            # shorter than the header
    s = np.zeros(20, dtype=np.uint8)
    s[8] = 100                                                # n = 100 rows, no body
    s[16] = 15
    s[12] = 15
    with pytest.raises(buff_amd.BuffError):
        gpu_ctx.load(s, 10 ** 4)
    # semantic mode encodes a constant column
    ch = gpu_ctx.encode(np.full(100, 2.5), 10 ** 2, flags=buff_amd.BUFF_SEMANTIC)
    assert np.all(ch.materialize() == 2.5)
    w, c = ch.filter("equal", 2.5)
    assert c == 100
    ch.free()


def test_hbm_budget_admission():
    import buff_amd

    ctx = buff_amd.Context(0, hbm_budget=1 << 20)
    x = dg.random_walk(1 << 20, 4, 5)            # 8 MiB of input alone exceeds 1 MiB
    with pytest.raises(buff_amd.OutOfMemory):
        ctx.encode(x, 10 ** 4)
    ctx.destroy()


def test_empty_column(gpu_ctx, oracle):
    ref = oracle.encode(np.zeros(0), 10 ** 4)
    ch = gpu_ctx.encode(np.zeros(0), 10 ** 4)
    assert np.array_equal(ch.serialize(), ref)
    assert ch.materialize().size == 0
    w, c = ch.filter("range", 1.0)
    assert c == 0
    assert ch.sum()[0] == oracle.byte_fixed_sum(ref, 10 ** 4)[0]
    ch.free()
