"""Batched execution on the GPU through the C ABI.

A batch of chunks with uneven lengths and mixed precision (the batched
workload's shape, scaled down, plus majority-value columns whose rows tie on
their top planes) is encoded with the one-launch batch encoder;
every chunk must serialize to the oracle's byte_fixed stream. Then mixed query
batches (range / equality filters, sums, maxima, materializations at full and
reduced precision, compat and semantic) run as one graph launch each, and every
result must equal the oracle's answer on the chunk's reference stream (the
closed forms for BUFF_SEMANTIC) and the single-chunk call on the same chunk:
bitmaps, counts, the reference f64 sum bit for bit, exact sums, max offsets and
values, materialized doubles bit for bit, and the same error status. Graph
replays with different batches, and the autotuner's choice, must not change any
answer. The 4096-chunk batch at its real size is in test_gpu_realsize.py.
"""
import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg

pytestmark = pytest.mark.gpu


def _columns(seed=5, nchunks=48):
    rng = np.random.default_rng(seed)
    cols = []
    for i in range(nchunks):
        n = int(np.exp(rng.uniform(np.log(1024), np.log(300_000))))
        p = int(rng.integers(1, 7))
        cols.append((dg.random_walk(n, p, seed + 100 + i), 10 ** p))
    # edge chunks: unique maximum in row 0, a tail-only width, F % 8 == 0, one row, ragged
    cols.append((np.concatenate([[50.0], np.round(np.linspace(0, 10, 70_000), 2)]), 10 ** 2))
    cols.append((np.round(rng.integers(0, 5, 3001) * 0.1, 1), 10))
    cols.append((np.round(np.linspace(0, 7.9, 4097), 1), 10))
    cols.append((np.array([3.25]), 10 ** 2))
    cols.append((dg.random_walk(1025, 3, 77), 10 ** 3))
    # majority-value columns: rows tie on the top planes in nearly every tile
    cols.append((dg.outlier_heavy(150_000, 3, 901), 10 ** 3))
    cols.append((dg.outlier_heavy(33_333, 3, 902), 10 ** 3))
    return cols


@pytest.fixture(scope="module")
def batch(gpu_ctx):
    cols = _columns()
    x = np.concatenate([c[0] for c in cols])
    row_off = np.concatenate([[0], np.cumsum([len(c[0]) for c in cols])]).astype(np.uint64)
    scales = np.array([c[1] for c in cols], dtype=np.uint64)
    xd = gpu_ctx.alloc(x.nbytes)
    xd.upload(x)
    b = gpu_ctx.batch_encode_device(xd.ptr, row_off, scales)
    yield b, cols
    b.free()
    xd.free()


def test_batch_encode_bit_exact(batch, oracle):
    b, cols = batch
    for i, ((x, scale), ch) in enumerate(zip(cols, b.chunks)):
        ref = oracle.encode(x, scale)
        got = ch.serialize()
        assert np.array_equal(got, ref), (i, len(x), scale)


def _mix(cols, seed, nq=600):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(nq):
        k = int(rng.integers(0, len(cols)))
        x, scale = cols[k]
        op = int(rng.integers(0, 5))
        pred = float(x[int(rng.integers(0, len(x)))]) if rng.random() < 0.7 else float(
            np.quantile(x, rng.random(), method="lower") + rng.choice([-1.0, 0.0, 1e-3]))
        prec = -1
        if op == N.BUFF_Q_MAT:
            prec = int(rng.choice([-1, 1, 2, 3]))
        flags = N.BUFF_SEMANTIC if rng.random() < 0.3 else 0
        qs.append((op, k, pred, prec, flags))
    qs.append((N.BUFF_Q_SUM, len(cols) + 5, 0.0, -1, 0))   # outside the batch
    return qs


def _check(b, cols, qs, res):
    for (op, k, pred, prec, flags), r in zip(qs, res):
        if k >= len(cols):
            assert r.status == N.BUFF_E_INVALID
            continue
        ch = b.chunks[k]
        n = len(cols[k][0])
        if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
            w, c = ch.filter("range" if op == N.BUFF_Q_GT else "equal", pred, flags=flags)
            assert r.status == 0 and r.count == c, (op, k, pred)
            assert np.array_equal(b.bitmap(r, n), w), (op, k, pred)
        elif op == N.BUFF_Q_SUM:
            s, e, _ = ch.sum()
            assert r.status == 0
            assert (r.value == s or (np.isnan(s) and np.isnan(r.value))) and r.exact == e, (k, r.value, s)
        elif op == N.BUFF_Q_MAX:
            v, off = ch.max()
            assert r.status == 0 and r.max_off == off and r.value == v, (k, r.max_off, off)
        else:
            try:
                want = ch.materialize(prec, flags)
            except N.BuffError as err:
                assert r.status == err.code, (k, prec, flags, r.status, err.code)
                continue
            assert r.status == 0, (k, prec, flags, r.status)
            got = b.values(r, n)
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), (k, prec, flags)


def _check_oracle(b, cols, streams, qs, res, oracle):
    """Every batched result against the oracle on the chunk's reference stream
    (BUFF_SEMANTIC answers against the closed forms)."""
    from tests import closed_form as cf

    for (op, k, pred, prec, flags), r in zip(qs, res):
        if k >= len(cols):
            continue
        s, scale = streams[k], cols[k][1]
        n = len(cols[k][0])
        if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
            kind = "range" if op == N.BUFF_Q_GT else "equal"
            if flags & N.BUFF_SEMANTIC:
                m = cf.filter_bits(s, scale, kind, pred, semantic=True)
                w_ref, c_ref = cf.to_words(m), int(m.sum())
            else:
                w_ref, c_ref, _ = oracle.buff_filter(s, scale, kind, pred)
            assert r.status == 0 and r.count == c_ref, ("oracle", op, k, pred)
            assert np.array_equal(b.bitmap(r, n), w_ref), ("oracle", op, k, pred)
        elif op == N.BUFF_Q_SUM:
            s_ref, _ = oracle.byte_fixed_sum(s, scale)
            tot, _ = cf.exact_sum(s)
            assert r.status == 0 and (r.value == s_ref or (np.isnan(s_ref) and np.isnan(r.value))), ("oracle", k)
            assert abs(r.exact - float(tot)) <= 1e-12 * max(1.0, abs(float(tot))), ("oracle", k)
        elif op == N.BUFF_Q_MAX:
            m_ref, off_ref, _ = oracle.byte_fixed_max(s, scale)
            assert r.status == 0 and r.max_off == off_ref and r.value == m_ref, ("oracle", k)
        else:
            if flags & N.BUFF_SEMANTIC:
                try:
                    want = cf.decode(s) if prec < 0 else cf.decode_semantic_precision(s, prec)
                except (ValueError, KeyError):
                    want = cf.decode(s)   # more decimals than stored: every bit
            else:
                try:
                    want = oracle.decode(s, scale) if prec < 0 else oracle.decode_with_precision(s, scale, prec)
                except oracle.RefPanic:
                    assert r.status == N.BUFF_E_REF_PANIC, ("oracle", k, prec)
                    continue
                except oracle.RefUnsupported:
                    assert r.status == N.BUFF_E_UNSUPPORTED, ("oracle", k, prec)
                    continue
                h = cf.parse(s)
                if prec >= 0 and cf.precision_R(h, prec) // 8 > h["P"]:
                    assert r.status == N.BUFF_E_UNSUPPORTED, ("oracle", k, prec)
                    continue
            assert r.status == 0, ("oracle", k, prec, flags, r.status)
            assert np.array_equal(b.values(r, n).view(np.uint64), want.view(np.uint64)), ("oracle", k, prec, flags)


def test_batch_queries_match_single_calls(batch, oracle):
    b, cols = batch
    streams = [oracle.encode(x, scale) for x, scale in cols]
    for seed in (1, 2):
        qs = _mix(cols, seed)
        print("batch", seed, len(qs), flush=True)
        res = b.query(qs)
        _check(b, cols, qs, res)
        _check_oracle(b, cols, streams, qs, res, oracle)


def test_graph_replay_and_tuning(batch):
    b, cols = batch
    qs1, qs2 = _mix(cols, 11, 300), _mix(cols, 12, 900)
    r1 = b.query(qs1)
    r2 = b.query(qs2)
    _check(b, cols, qs2, r2)   # the arenas hold the latest batch's bitmaps and values
    r1b = b.query(qs1)
    assert [(r.status, r.count, r.value, r.exact, r.max_off) for r in r1] == \
           [(r.status, r.count, r.value, r.exact, r.max_off) for r in r1b]
    grid, tiles, first = b.tune(_mix(cols, 13, 200))
    assert grid in (1024, 2048, 4096, 8192) and tiles in (16, 64, 256) and first in (1, 8)
    _check(b, cols, qs1, b.query(qs1))


def test_batch_encode_rejects_panicking_chunk(gpu_ctx):
    x = np.concatenate([dg.random_walk(5000, 2, 1), np.full(3000, 7.25)])
    xd = gpu_ctx.alloc(x.nbytes)
    xd.upload(x)
    try:
        with pytest.raises(N.RefPanic):
            gpu_ctx.batch_encode_device(xd.ptr, [0, 5000, 8000], [100, 100])
    finally:
        xd.free()
