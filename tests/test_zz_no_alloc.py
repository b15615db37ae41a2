"""buff_sum_noalloc (the allocation-free sum) against buff_sum, through the C ABI.

Two calls in a row, the second on a smaller column (the static scratch keeps the
first call's contents), bit for bit: the reference f64 sum, the exact sum and
the nine per-sub-column sums; the context's HBM use does not change.
"""
import numpy as np
import pytest

import buff_amd
from buff_amd import datagen as dg

pytestmark = pytest.mark.gpu


def test_sum_noalloc_matches_buff_sum():
    ctx = buff_amd.Context(0)
    cols = [(dg.random_walk(3_000_001, 4, 71), 10 ** 4),     # larger first
            (dg.uniform(100_003, -1000.0, 1000.0, 2, 72), 10 ** 2),
            (np.round(np.linspace(100, 107.9, 4097), 1), 10)]   # F = 8: no tail, sum without the base term
    chunks = [ctx.encode(x, s) for x, s in cols]
    try:
        for ch in chunks:
            want = ch.sum()
            ctx.synchronize()
            used = ctx.hbm_used()
            got = ch.sum_noalloc()
            assert ctx.hbm_used() == used
            assert np.float64(got[0]).view(np.uint64) == np.float64(want[0]).view(np.uint64)
            assert np.float64(got[1]).view(np.uint64) == np.float64(want[1]).view(np.uint64)
            assert np.array_equal(got[2], want[2])
        # repeated on the first (largest) column after the smaller ones
        assert chunks[0].sum_noalloc()[0] == chunks[0].sum()[0]
        maj = ctx.encode(dg.outlier_heavy(5000, 3, 73), 10 ** 3, flags=buff_amd.BUFF_MAJORITY)
        with pytest.raises(buff_amd.Unsupported):
            maj.sum_noalloc()
        maj.free()
    finally:
        for ch in chunks:
            ch.free()
