"""Device-memory helpers and the one-launch reductions' slot state.

* buff_dev_copy (the streaming-copy ceiling kernel) copies exactly the bytes
  asked for, at several sizes, and rejects sizes or addresses that are not
  multiples of 16 and overlapping ranges; buff_dev_read (the streaming-read
  ceiling) leaves memory unchanged and rejects the same misuses.
* Sums, maxima and min/max passes fold into slot sets kept across calls (two
  alternating sets for sums and the min/max pass, per-call tags for the max):
  interleaved calls on different chunks, many times over, must each give the
  answer the chunk gives alone (checked against numpy on the stored offsets).
* Results reach the host through each context's mailbox: two contexts used
  from two threads at once, each with its own chunks, must each get their own
  sums, maxima and encodes, call after call.
* Encode takes float64 input at any 8-byte aligned device address: rows that
  start 8 bytes into a buffer must give the same chunk, byte for byte, as the
  same rows at a 256-byte aligned address.
"""
from fractions import Fraction

import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg
from tests import closed_form as cf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096 + 16, (1 << 20) + 48, 64 << 20])
def test_dev_copy_exact(gpu_ctx, nbytes):
    rng = np.random.default_rng(nbytes)
    src_h = rng.integers(0, 256, nbytes + 64, dtype=np.uint8)
    src = gpu_ctx.alloc(nbytes + 64)
    dst = gpu_ctx.alloc(nbytes + 64)
    try:
        src.upload(src_h)
        dst.fill(0xAB)
        dst.copy_from(src, nbytes)
        got = dst.download(np.uint8, nbytes + 64)
        assert np.array_equal(got[:nbytes], src_h[:nbytes])
        assert np.all(got[nbytes:] == 0xAB)   # nothing past the copy
    finally:
        src.free()
        dst.free()


def test_dev_read_runs_and_rejects(gpu_ctx):
    a = gpu_ctx.alloc(1 << 20)
    try:
        a.fill(7)
        a.stream_read()            # reads and discards; the buffer is unchanged
        a.stream_read(4096 + 16)
        assert np.all(a.download(np.uint8, 1 << 20) == 7)
        with pytest.raises(N.BuffError):
            N.check(N.lib().buff_dev_read(gpu_ctx.handle, a.ptr, 100))       # not a multiple of 16
        with pytest.raises(N.BuffError):
            N.check(N.lib().buff_dev_read(gpu_ctx.handle, a.ptr + 8, 1024))  # unaligned address
    finally:
        a.free()


def test_dev_copy_rejects(gpu_ctx):
    a = gpu_ctx.alloc(4096)
    b = gpu_ctx.alloc(4096)
    try:
        with pytest.raises(N.BuffError):
            N.check(N.lib().buff_dev_copy(gpu_ctx.handle, b.ptr, a.ptr, 100))   # not a multiple of 16
        with pytest.raises(N.BuffError):
            N.check(N.lib().buff_dev_copy(gpu_ctx.handle, b.ptr + 8, a.ptr, 1024))   # unaligned address
        with pytest.raises(N.BuffError):
            N.check(N.lib().buff_dev_copy(gpu_ctx.handle, a.ptr + 1024, a.ptr, 2048))   # overlapping
    finally:
        a.free()
        b.free()


def test_slot_reductions_interleaved(gpu_ctx):
    cols = [(dg.random_walk(300_000, 4, 1), 10 ** 4), (dg.uniform(70_001, -1000, 1000, 2, 2), 100),
            (dg.random_walk(2_000_003, 2, 3), 100), (np.round(np.linspace(0, 7.9, 5000), 1), 10)]
    chunks = [gpu_ctx.encode(x, s) for x, s in cols]
    try:
        expect = []   # exact sum and max offset of each chunk, from its stored offsets
        for (x, s), ch in zip(cols, chunks):
            inf = ch.info()
            off = (cf.fixed(x, inf.dlen) - inf.base).astype(np.uint64)   # every offset < 2^41: the u64 sum is exact
            exact = float(Fraction(inf.n * inf.base + int(off.sum(dtype=np.uint64)), 1 << inf.dlen))
            expect.append((exact, int(off.max())))
        rng = np.random.default_rng(9)
        for step in range(60):
            k = int(rng.integers(0, len(chunks)))
            ch = chunks[k]
            if step % 2 == 0:
                _, ex, _ = ch.sum()
                assert ex == expect[k][0], (step, k, ex, expect[k][0])
            else:
                _, off = ch.max()
                assert off == expect[k][1], (step, k, off, expect[k][1])
            if step % 7 == 3:   # an encode between the reductions runs the min/max pass again
                x, s = cols[k]
                again = gpu_ctx.encode(x, s)
                try:
                    assert np.array_equal(again.serialize(), ch.serialize()), (step, k)
                finally:
                    again.free()
    finally:
        for ch in chunks:
            ch.free()


@pytest.mark.parametrize("n", [1024, 5 * 1024 + 3, 300_001])
def test_encode_aligned_and_unaligned_input(gpu_ctx, n):
    x = dg.random_walk(n, 4, 21)
    a = gpu_ctx.alloc((n + 2) * 8)
    try:
        a.upload(np.concatenate([x, [0.0, 0.0]]))
        ch0 = gpu_ctx.encode_device(a.ptr, n, 10 ** 4)   # 256-byte aligned
        try:
            want = ch0.serialize()
            a.upload(np.concatenate([[0.0], x, [0.0]]))
            ch1 = gpu_ctx.encode_device(a.ptr + 8, n, 10 ** 4)   # 8 bytes in: not 16-byte aligned
            try:
                assert np.array_equal(ch1.serialize(), want), n
            finally:
                ch1.free()
            host = gpu_ctx.encode(x, 10 ** 4)
            try:
                assert np.array_equal(host.serialize(), want), n
            finally:
                host.free()
        finally:
            ch0.free()
    finally:
        a.free()


def test_mailboxes_of_two_contexts_in_two_threads(gpu_ctx):
    import threading

    import buff_amd

    def work(ctx, seed, errors):
        try:
            x = dg.random_walk(500_000 + seed, 3, seed)
            ch = ctx.encode(x, 1000)
            try:
                inf = ch.info()
                off = (cf.fixed(x, inf.dlen) - inf.base).astype(np.uint64)
                exact = float(Fraction(inf.n * inf.base + int(off.sum(dtype=np.uint64)), 1 << inf.dlen))
                want = ch.serialize()
                for i in range(40):
                    assert ch.sum()[1] == exact, (seed, i)
                    assert ch.max()[1] == int(off.max()), (seed, i)
                    if i % 10 == 0:
                        again = ctx.encode(x, 1000)
                        try:
                            assert np.array_equal(again.serialize(), want), (seed, i)
                        finally:
                            again.free()
            finally:
                ch.free()
        except BaseException as e:   # reported by the main thread
            errors.append(e)

    other = buff_amd.Context(0)
    try:
        errors = []
        ts = [threading.Thread(target=work, args=(c, seed, errors)) for c, seed in ((gpu_ctx, 11), (other, 12))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errors, errors
    finally:
        other.destroy()
