"""Deterministic synthetic columns of the benchmark workloads.

All generators take an explicit seed and use numpy's PCG64 stream, so the same
call yields the same doubles on every machine with the same numpy.
  * random_walk: x_0 = 0, x_{t+1} = x_t + N(0, 1), each stored value rounded
    to `decimals` (the walk accumulates the unrounded values).
  * uniform: U[lo, hi] rounded to `decimals`.
  * outlier_heavy: 95% N(100, 0.5), 5% +-10^U[6, 9], rounded to `decimals`.
  * positive_spikes / wide_spikes: N(100, 0.5) with a fraction of rows
    100 + 10^U[3, 7] / 100 + 10^U[5, 9] (the residue format's outlier columns).
Rounding is numpy.round (round half to even on x * 10^d).
"""
import numpy as np


def random_walk(n, decimals, seed):
    rng = np.random.default_rng(seed)
    steps = rng.standard_normal(n)
    steps[0] = 0.0
    walk = np.cumsum(steps)
    return np.round(walk, decimals)


def uniform(n, lo, hi, decimals, seed):
    rng = np.random.default_rng(seed)
    return np.round(rng.uniform(lo, hi, n), decimals)


def outlier_heavy(n, decimals, seed, frac=0.05):
    rng = np.random.default_rng(seed)
    x = rng.normal(100.0, 0.5, n)
    out = rng.random(n) < frac
    k = int(out.sum())
    mag = 10.0 ** rng.uniform(6.0, 9.0, k)
    sign = np.where(rng.random(k) < 0.5, -1.0, 1.0)
    x[out] = sign * mag
    return np.round(x, decimals)


def positive_spikes(n, decimals, seed, frac=0.05):
    """95% N(100, 0.5) and `frac` of rows 100 + 10^U[3,7]: offsets from the
    minimum are small except the spikes', so the top byte columns are mostly
    zero (the residue format's outlier columns)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(100.0, 0.5, n)
    out = rng.random(n) < frac
    x[out] = 100.0 + 10.0 ** rng.uniform(3.0, 7.0, int(out.sum()))
    return np.round(x, decimals)


def wide_spikes(n, decimals, seed, frac=0.03):
    """95%+ N(100, 0.5) and `frac` of rows 100 + 10^U[5,9]: offsets up to 2^30
    units, so three leading byte columns are mostly zero."""
    rng = np.random.default_rng(seed)
    x = rng.normal(100.0, 0.5, n)
    out = rng.random(n) < frac
    x[out] = 100.0 + 10.0 ** rng.uniform(5.0, 9.0, int(out.sum()))
    return np.round(x, decimals)


def percentile_pred(x, frac_above):
    """A predicate with about `frac_above` of the rows strictly above it."""
    q = np.quantile(x, 1.0 - frac_above, method="lower")
    return float(q)
