"""The serving loop on the GPU: continuous batching must not change any answer.

Resident chunks of outlier-heavy data (the served workload's distribution, at
test sizes) take a request mix of 60% range filters, 25% sums and 15%
materializations at 1-3 decimals. Every request's result is computed alone
first (single-chunk calls), then all requests are served concurrently from
several client threads and as an open-loop trace; each served result must be
bit-identical to the alone result (bitmaps, counts, reference and exact sums,
materialized doubles). Stats show requests were coalesced into batches, and a
small output arena forces requests to wait for memory without losing any.
"""
import threading

import numpy as np
import pytest

import buff_amd
from buff_amd import _native as N
from buff_amd import datagen as dg

pytestmark = pytest.mark.gpu


def _requests(lens, seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(0, len(lens)))
        u = rng.random()
        if u < 0.6:
            out.append((N.BUFF_Q_GT, k, float(rng.choice([99.2, 99.8, 100.0, 100.4, 101.0, -5e6, 3e8])), -1, 0))
        elif u < 0.85:
            out.append((N.BUFF_Q_SUM, k, 0.0, -1, 0))
        else:
            out.append((N.BUFF_Q_MAT, k, 0.0, int(rng.integers(1, 4)), N.BUFF_SEMANTIC))
    return out


@pytest.fixture(scope="module")
def served(gpu_ctx):
    rng = np.random.default_rng(3)
    lens = [int(v) for v in np.exp(rng.uniform(np.log(2000), np.log(300_000), 20))]
    xs = [dg.outlier_heavy(n, 3, 40 + i) for i, n in enumerate(lens)]
    chunks = [gpu_ctx.encode(x, 10 ** 3) for x in xs]
    yield chunks, lens, xs
    for c in chunks:
        c.free()


def _alone(chunks, q):
    op, k, pred, prec, flags = q
    ch = chunks[k]
    if op == N.BUFF_Q_GT:
        return ch.filter("range", pred)
    if op == N.BUFF_Q_SUM:
        return ch.sum()[:2]
    return ch.materialize(prec, flags)


def _outs(ctx, lens, reqs):
    bufs = []
    for op, k, _, _, _ in reqs:
        if op == N.BUFF_Q_GT:
            bufs.append(ctx.alloc(((lens[k] + 63) // 64) * 8))
        elif op == N.BUFF_Q_MAT:
            bufs.append(ctx.alloc(lens[k] * 8))
        else:
            bufs.append(None)
    return bufs


def _same(ctx, lens, q, want, r, buf):
    op, k = q[0], q[1]
    assert r.status == 0, (q, r.status)
    if op == N.BUFF_Q_GT:
        words = buf.download(np.uint32, ((lens[k] + 63) // 64) * 2)[: (lens[k] + 31) // 32]
        assert r.count == want[1] and np.array_equal(words, want[0]), q
    elif op == N.BUFF_Q_SUM:
        assert (r.value, r.exact) == want, q
    else:
        vals = buf.download(np.float64, lens[k])
        assert np.array_equal(vals.view(np.int64), want.view(np.int64)), q


def _oracle_same(oracle, streams, lens, q, r, buf):
    """A served result against the oracle on the chunk's reference stream (the
    semantic reduced-precision decode against the closed form)."""
    from tests import closed_form as cf

    op, k, pred, prec, flags = q
    s = streams[k]
    assert r.status == 0, (q, r.status)
    if op == N.BUFF_Q_GT:
        w_ref, c_ref, _ = oracle.buff_filter(s, 10 ** 3, "range", pred)
        assert r.count == c_ref, ("oracle", q)
        if buf is not None:
            words = buf.download(np.uint32, ((lens[k] + 63) // 64) * 2)[: (lens[k] + 31) // 32]
            assert np.array_equal(words, w_ref), ("oracle", q)
    elif op == N.BUFF_Q_SUM:
        s_ref, _ = oracle.byte_fixed_sum(s, 10 ** 3)
        tot, _ = cf.exact_sum(s)
        assert r.value == s_ref and abs(r.exact - float(tot)) <= 1e-12 * abs(float(tot)), ("oracle", q)
    elif buf is not None:
        want = cf.decode_semantic_precision(s, prec)
        vals = buf.download(np.float64, lens[k])
        assert np.array_equal(vals.view(np.uint64), want.view(np.uint64)), ("oracle", q)


def test_served_equals_alone_threads(gpu_ctx, served, oracle):
    chunks, lens, xs = served
    reqs = _requests(lens, 1, 400)
    alone = [_alone(chunks, q) for q in reqs]
    bufs = _outs(gpu_ctx, lens, reqs)
    srv = buff_amd.Server(gpu_ctx, max_batch=64, slots=2)
    try:
        ids = [srv.add_chunk(c) for c in chunks]
        assert ids == list(range(len(chunks)))
        results = [None] * len(reqs)

        def client(part):
            tickets = []
            for i in part:
                op, k, pred, prec, flags = reqs[i]
                tickets.append((i, srv.submit(op, k, pred, prec, flags, bufs[i].ptr if bufs[i] else None)))
            for i, t in tickets:
                results[i] = srv.wait(t)[0]

        th = [threading.Thread(target=client, args=(range(j, len(reqs), 4),)) for j in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        batches, served_n, largest = srv.stats()
        assert served_n == len(reqs) and batches < len(reqs) and largest > 1
        streams = [oracle.encode(x, 10 ** 3) for x in xs]
        for q, want, r, b in zip(reqs, alone, results, bufs):
            _same(gpu_ctx, lens, q, want, r, b)
            _oracle_same(oracle, streams, lens, q, r, b)
    finally:
        srv.destroy()
        for b in bufs:
            if b:
                b.free()


def test_served_trace_and_memory_admission(gpu_ctx, served, oracle):
    chunks, lens, xs = served
    reqs = _requests(lens, 2, 300)
    alone = [_alone(chunks, q) for q in reqs]
    bufs = _outs(gpu_ctx, lens, reqs)
    # arena of 4 MiB: requests without an output address wait for memory to free up
    srv = buff_amd.Server(gpu_ctx, max_batch=256, slots=2, arena_bytes=4 << 20)
    try:
        for c in chunks:
            srv.add_chunk(c)
        arrival = np.cumsum(np.random.default_rng(9).exponential(1 / 20000.0, len(reqs)))
        res, lat = srv.run_trace(reqs, arrival, [b.ptr if b else 0 for b in bufs])
        assert np.all(lat > 0)
        streams = [oracle.encode(x, 10 ** 3) for x in xs]
        for q, want, r, b in zip(reqs, alone, res, bufs):
            _same(gpu_ctx, lens, q, want, r, b)
            _oracle_same(oracle, streams, lens, q, r, b)
        # the same requests without output buffers: only counts and sums come back
        res2, _ = srv.run_trace(reqs, arrival)
        for q, want, r in zip(reqs, alone, res2):
            assert r.status == 0
            if q[0] == N.BUFF_Q_GT:
                assert r.count == want[1]
            elif q[0] == N.BUFF_Q_SUM:
                assert (r.value, r.exact) == want
        batches, served_n, largest = srv.stats()
        assert served_n == 2 * len(reqs)
    finally:
        srv.destroy()
        for b in bufs:
            if b:
                b.free()


def _d2h(ctx, addr, dtype, count):
    out = np.zeros(count, dtype=dtype)
    N.check(N.lib().buff_memcpy_d2h(ctx.handle, out.ctypes.data, addr, out.nbytes))
    return out


def test_served_arena_outputs_held_until_release(gpu_ctx, served, oracle):
    # admission by memory: an arena with room for one bitmap; a served bitmap stays
    # readable until it is released, and the next request waits for that release
    chunks, lens, xs = served
    k = int(np.argmax(lens))
    n = lens[k]
    nb = ((n + 63) // 64) * 8
    arena = (nb + 255) // 256 * 256
    stream = oracle.encode(xs[k], 10 ** 3)
    srv = buff_amd.Server(gpu_ctx, max_batch=64, slots=2, arena_bytes=arena)
    try:
        for c in chunks:
            srv.add_chunk(c)
        t1 = srv.submit(N.BUFF_Q_GT, k, 100.0)
        r1, _ = srv.wait(t1)
        assert r1.status == 0 and r1.out_offset
        w_ref, c_ref, _ = oracle.buff_filter(stream, 10 ** 3, "range", 100.0)
        assert r1.count == c_ref
        t2 = srv.submit(N.BUFF_Q_GT, k, 100.4)
        t3 = srv.submit(N.BUFF_Q_SUM, k)   # behind t2 in arrival order
        import time
        time.sleep(0.2)
        assert not srv.poll(t2) and not srv.poll(t3), "admitted without arena memory"
        assert srv.arena_stats()[0] == arena
        got = _d2h(gpu_ctx, r1.out_offset, np.uint32, (n + 31) // 32)
        assert np.array_equal(got, w_ref), "a held bitmap was overwritten"
        srv.release(r1.out_offset)
        r2, _ = srv.wait(t2)
        r3, _ = srv.wait(t3)
        assert r2.status == 0 and r3.status == 0 and r3.out_offset == 0
        w2, c2, _ = oracle.buff_filter(stream, 10 ** 3, "range", 100.4)
        assert r2.count == c2 and np.array_equal(_d2h(gpu_ctx, r2.out_offset, np.uint32, (n + 31) // 32), w2)
        srv.release(r2.out_offset)
        with pytest.raises(N.BuffError):
            srv.release(r2.out_offset)   # already returned
        # an output larger than the whole arena can never be admitted
        t4 = srv.submit(N.BUFF_Q_MAT, k, 0.0, 2, N.BUFF_SEMANTIC)
        r4, _ = srv.wait(t4)
        assert r4.status == N.BUFF_E_OOM and r4.out_offset == 0
        in_use, peak, cap = srv.arena_stats()
        assert in_use == 0 and peak == arena and cap == arena
    finally:
        srv.destroy()


def test_trace_per_call_baseline(gpu_ctx, served, oracle):
    # the plain one-call-per-request replay (the baseline continuous batching is
    # measured against) answers like the single-chunk calls
    chunks, lens, xs = served
    reqs = _requests(lens, 3, 120)
    bufs = _outs(gpu_ctx, lens, reqs)
    try:
        arrival = np.cumsum(np.random.default_rng(4).exponential(1 / 5000.0, len(reqs)))
        res, lat = gpu_ctx.trace_per_call(chunks, reqs, arrival, [b.ptr if b else 0 for b in bufs])
        assert np.all(lat > 0)
        streams = [oracle.encode(x, 10 ** 3) for x in xs]
        for q, r, b in zip(reqs, res, bufs):
            assert r.status == 0
            if q[0] == N.BUFF_Q_GT:
                w_ref, _, _ = oracle.buff_filter(streams[q[1]], 10 ** 3, "range", q[2])
                words = b.download(np.uint32, ((lens[q[1]] + 63) // 64) * 2)[: (lens[q[1]] + 31) // 32]
                assert np.array_equal(words, w_ref), q
            elif q[0] == N.BUFF_Q_SUM:
                assert (r.value, r.exact) == tuple(chunks[q[1]].sum()[:2]), q
            else:
                from tests import closed_form as cf
                want = cf.decode_semantic_precision(streams[q[1]], q[3])
                vals = b.download(np.float64, lens[q[1]])
                assert np.array_equal(vals.view(np.uint64), want.view(np.uint64)), q
    finally:
        for b in bufs:
            if b:
                b.free()


# ---- buff-major chunks (the served workload's layout) ---------------------------

@pytest.fixture(scope="module")
def served_major(gpu_ctx, oracle):
    rng = np.random.default_rng(8)
    lens = [int(v) for v in np.exp(rng.uniform(np.log(5000), np.log(300_000), 10))]
    xs = [dg.outlier_heavy(n, 3, 70 + i) for i, n in enumerate(lens)]
    xs.append(dg.random_walk(40_000, 3, 99))   # no majority byte: a buff-major stream without sparse columns
    lens.append(40_000)
    chunks = [gpu_ctx.encode(x, 10 ** 3, flags=buff_amd.BUFF_MAJORITY) for x in xs]
    streams = [c.serialize() for c in chunks]
    for x, s in zip(xs, streams):
        assert np.array_equal(s, oracle.major_encode(x, 10 ** 3))
    yield chunks, lens, xs, streams
    for c in chunks:
        c.free()


def _major_requests(lens, seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(0, len(lens)))
        u = rng.random()
        sem = N.BUFF_SEMANTIC if rng.random() < 0.4 else 0
        pred = float(rng.choice([99.2, 99.8, 100.0, 100.3, 100.4, 101.0, -5e6, 3e8, 1e10]))
        if u < 0.45:
            out.append((N.BUFF_Q_GT, k, pred, -1, sem))
        elif u < 0.55:
            out.append((N.BUFF_Q_EQ, k, pred, -1, sem))
        elif u < 0.8:
            out.append((N.BUFF_Q_SUM, k, 0.0, -1, 0))
        elif u < 0.95:
            out.append((N.BUFF_Q_MAT, k, 0.0, int(rng.integers(1, 4)), N.BUFF_SEMANTIC))
        else:
            out.append((N.BUFF_Q_MAT, k, 0.0, -1, sem))
    return out


def _major_want(oracle, x, s, q):
    """The oracle's answer (compat) or numpy's on the true offsets (BUFF_SEMANTIC)."""
    from tests import closed_form as cf

    op, k, pred, prec, flags = q
    hdr = np.frombuffer(np.asarray(s[:20], dtype=np.uint8).tobytes(), dtype=np.uint32)
    h = {"ilen": int(hdr[3]), "dlen": int(hdr[4])}   # header fields (the body is the buff-major layout)
    dlen = h["dlen"]
    F = (h["ilen"] + dlen) & 0xFFFFFFFF
    f = cf.fixed(x, dlen)
    base = int(f.min())
    if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
        kind = "range" if op == N.BUFF_Q_GT else "equal"
        if flags & N.BUFF_SEMANTIC:
            fp = cf.fixed_pred(pred, 10 ** 3, dlen)
            m = np.full(len(x), kind == "range") if fp < base else (f > fp if kind == "range" else f == fp)
            return cf.to_words(m), int(m.sum())
        w, c = oracle.major_filter(s, 10 ** 3, kind, pred)[:2]
        return w, c
    if op == N.BUFF_Q_SUM:
        return oracle.major_sum(s, 10 ** 3)[0], sum(int(v) for v in f.astype(object))
    if prec < 0 and not flags & N.BUFF_SEMANTIC:
        return oracle.major_decode(s, 10 ** 3)
    try:
        R = F if prec < 0 else min(cf.precision_R(h, prec), F)
    except (ValueError, KeyError):
        R = F   # more decimals than stored: every bit
    sh = F - R
    off = (f - base).astype(np.uint64)
    return ((((off >> np.uint64(sh)) << np.uint64(sh)).astype(np.int64) + base).astype(np.float64) / 2.0 ** dlen)


def test_served_buff_major_matches_oracle(gpu_ctx, served_major, oracle):
    # buff-major chunks served with continuous batching: every answer equals the
    # single-chunk call's and the oracle's (numpy's on the true offsets for
    # BUFF_SEMANTIC); a max on a chunk with sparse columns runs the single-chunk
    # stages when its batch completes and equals the oracle's too
    from fractions import Fraction

    chunks, lens, xs, streams = served_major
    reqs = _major_requests(lens, 21, 300)
    for k in range(len(chunks)):
        reqs.insert(37 * k + 5, (N.BUFF_Q_MAX, k, 0.0, -1, 0))
        reqs.append((N.BUFF_Q_MAX, k, 0.0, -1, N.BUFF_SEMANTIC))
    bufs = _outs(gpu_ctx, lens, [(N.BUFF_Q_GT if q[0] == N.BUFF_Q_EQ else q[0],) + q[1:] for q in reqs])
    srv = buff_amd.Server(gpu_ctx, max_batch=64, slots=2)
    try:
        for c in chunks:
            srv.add_chunk(c)
        res, _ = srv.run_trace(reqs, np.zeros(len(reqs)), [b.ptr if b else 0 for b in bufs])
        batches, _, largest = srv.stats()
        assert largest > 1
        for q, r, b in zip(reqs, res, bufs):
            op, k, pred, prec, flags = q
            n = lens[k]
            if op == N.BUFF_Q_MAX:
                assert r.status == 0, (q, r.status)
                alone = chunks[k].max_ex(flags=flags)
                assert (r.value, r.max_off) == alone[:2], (q, r.value, r.max_off, alone)
                if flags & N.BUFF_SEMANTIC:
                    from tests import closed_form as cf
                    f = cf.fixed(xs[k], int(chunks[k].info().dlen))
                    assert r.max_off == int(f.max() - f.min()), q
                else:
                    v_ref, off_ref, _ = oracle.major_max(streams[k], 10 ** 3)
                    assert r.value == v_ref and r.max_off == off_ref, (q, r.value, r.max_off, v_ref, off_ref)
                continue
            assert r.status == 0, (q, r.status)
            want = _major_want(oracle, xs[k], streams[k], q)
            if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
                words = b.download(np.uint32, ((n + 63) // 64) * 2)[: (n + 31) // 32]
                assert r.count == want[1] and np.array_equal(words, want[0]), q
                alone = chunks[k].filter("range" if op == N.BUFF_Q_GT else "equal", pred, flags=flags)
                assert np.array_equal(alone[0], words), q
            elif op == N.BUFF_Q_SUM:
                ref, tot = want
                dlen = chunks[k].info().dlen
                assert r.value == ref, (q, r.value, ref)
                assert r.exact == float(Fraction(tot, 1 << int(dlen))), q
            else:
                got = b.download(np.float64, n)[: len(want)]
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), q
                alone = chunks[k].materialize(prec, flags)
                assert np.array_equal(alone.view(np.uint64), want.view(np.uint64)), q
    finally:
        srv.destroy()
        for b in bufs:
            if b:
                b.free()


def test_buff_major_semantic_reduced_precision(gpu_ctx, served_major, oracle):
    # single-chunk materialization of a buff-major chunk at 1..3 decimals
    # (BUFF_SEMANTIC: the top R stored bits); the reference has no such decode
    chunks, lens, xs, streams = served_major
    for k in (0, 3, len(chunks) - 1):
        for p in (0, 1, 2, 3, 5):
            want = _major_want(oracle, xs[k], streams[k], (N.BUFF_Q_MAT, k, 0.0, p, N.BUFF_SEMANTIC))
            got = chunks[k].materialize(p, flags=N.BUFF_SEMANTIC)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (k, p)
        with pytest.raises(N.Unsupported):
            chunks[k].materialize(2)


def test_served_residue_chunks_equal_single_chunk_calls(gpu_ctx, oracle):
    # residue chunks have no batched kernels: the server answers their requests
    # with the single-chunk calls when the batch completes (the reference's
    # answers, or the intended ones with BUFF_SEMANTIC; a filter or max without
    # BUFF_SEMANTIC is BUFF_E_UNSUPPORTED as alone), next to plain chunks
    # that are batched
    from tests.test_rapg_oracle import CASES

    names = ["spikes_5pct", "walk", "outlier_heavy", "spikes_p5"]
    xs = [CASES[m][0] for m in names] + [dg.random_walk(200_000, 3, 5)]
    scales = [CASES[m][1] for m in names] + [1000]
    flags = [N.BUFF_RESIDUE] * len(names) + [0]
    chunks = [gpu_ctx.encode(x, s, flags=f) for x, s, f in zip(xs, scales, flags)]
    lens = [len(x) for x in xs]
    streams = [oracle.residue_encode(x, s)[0] for x, s in zip(xs[:-1], scales[:-1])]
    sem = N.BUFF_SEMANTIC
    reqs = []
    for k, x in enumerate(xs):
        hi = float(np.quantile(x, 0.9, method="lower"))
        reqs += [(N.BUFF_Q_SUM, k, 0.0, -1, 0), (N.BUFF_Q_SUM, k, 0.0, -1, sem), (N.BUFF_Q_MAX, k, 0.0, -1, sem),
                 (N.BUFF_Q_MAX, k, 0.0, -1, 0), (N.BUFF_Q_GT, k, hi, -1, sem), (N.BUFF_Q_GT, k, hi, -1, 0),
                 (N.BUFF_Q_EQ, k, float(x[len(x) // 3]), -1, sem), (N.BUFF_Q_MAT, k, 0.0, -1, 0)]
        reqs += [(N.BUFF_Q_MAT, k, 0.0, p, 0) for p in (0, 1, 2, 3)]
    order = np.random.default_rng(4).permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    bufs = []
    for op, k, _, _, _ in reqs:
        if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
            bufs.append(gpu_ctx.alloc(((lens[k] + 63) // 64) * 8))
        elif op == N.BUFF_Q_MAT:
            bufs.append(gpu_ctx.alloc(lens[k] * 8))
        else:
            bufs.append(None)

    def alone(q):
        """The single-chunk call's status and answer."""
        op, k, pred, prec, fl = q
        ch = chunks[k]
        try:
            if op == N.BUFF_Q_SUM:
                return 0, ch.sum(fl)[:2]
            if op == N.BUFF_Q_MAX:
                return 0, ch.max_ex(flags=fl)[:2]
            if op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
                return 0, ch.filter("range" if op == N.BUFF_Q_GT else "equal", pred, flags=fl)
            return 0, ch.materialize(prec, fl)
        except N.BuffError as e:
            return e.code, None

    srv = buff_amd.Server(gpu_ctx, max_batch=16, slots=2)
    try:
        ids = [srv.add_chunk(c) for c in chunks]
        assert ids == list(range(len(chunks)))
        # every request at t = 0, then arrivals spread over 20 ms (single-chunk
        # calls of one batch run while the other slot's batch is in flight)
        spread = np.sort(np.random.default_rng(6).uniform(0, 0.02, len(reqs)))
        for arrivals in (np.zeros(len(reqs)), spread):
            res, _ = srv.run_trace(reqs, arrivals, [b.ptr if b else 0 for b in bufs])
            for q, r, b in zip(reqs, res, bufs):
                op, k, pred, prec, fl = q
                n = lens[k]
                code, want = alone(q)
                assert r.status == code, (q, r.status, code)
                if code:
                    continue
                if op == N.BUFF_Q_SUM:
                    assert (r.value, r.exact) == want, (q, r.value, r.exact, want)
                    if k < len(streams) and not fl:
                        assert r.value == oracle.residue_sum(streams[k])[0], q
                elif op == N.BUFF_Q_MAX:
                    assert (r.value, r.max_off) == want, (q, r.value, r.max_off, want)
                elif op in (N.BUFF_Q_GT, N.BUFF_Q_EQ):
                    words = b.download(np.uint32, ((n + 63) // 64) * 2)[: (n + 31) // 32]
                    assert r.count == want[1] and np.array_equal(words, want[0]), q
                else:
                    got = b.download(np.float64, n)[: len(want)]
                    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), q
                    if k < len(streams) and prec < 0:
                        ref = oracle.residue_decode(streams[k], -1)
                        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), q
            for b in bufs:   # a stale output from the first pass cannot pass for the second's
                if b:
                    b.fill(0xFF)
    finally:
        srv.destroy()
        for c in chunks:
            c.free()
        for b in bufs:
            if b:
                b.free()


def test_served_residue_outputs_in_the_arena(gpu_ctx, oracle):
    # a residue request without an output address writes into the server's
    # arena like any other: its result names the range, which reads back as the
    # single-chunk answer and returns to the arena on release
    from tests.test_rapg_oracle import CASES

    x, scale = CASES["spikes_5pct"]
    n = len(x)
    ch = gpu_ctx.encode(x, scale, flags=N.BUFF_RESIDUE)
    srv = buff_amd.Server(gpu_ctx, max_batch=8, slots=2, arena_bytes=64 << 20)
    try:
        srv.add_chunk(ch)
        pred = float(np.quantile(x, 0.9, method="lower"))
        t1 = srv.submit(N.BUFF_Q_GT, 0, pred, flags=N.BUFF_SEMANTIC)
        t2 = srv.submit(N.BUFF_Q_MAT, 0, precision=-1)
        r1, _ = srv.wait(t1)
        r2, _ = srv.wait(t2)
        assert r1.status == 0 and r1.out_offset and r2.status == 0 and r2.out_offset
        words, count = ch.filter("range", pred, flags=N.BUFF_SEMANTIC)
        assert r1.count == count
        assert np.array_equal(_d2h(gpu_ctx, r1.out_offset, np.uint32, (n + 31) // 32), words)
        want = oracle.residue_decode(oracle.residue_encode(x, scale)[0], -1)
        got = _d2h(gpu_ctx, r2.out_offset, np.float64, n)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        assert srv.arena_stats()[0] > 0
        srv.release(r1.out_offset)
        srv.release(r2.out_offset)
        assert srv.arena_stats()[0] == 0
    finally:
        srv.destroy()
        ch.free()
