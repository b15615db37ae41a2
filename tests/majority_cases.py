"""Columns that exercise the buff-major (majority / outlier) layout.

Built from chosen fixed-point offsets: with scale 1000 (dlen 11) the value
(base + off) / 2^11 is exact in float64, so fetch_fixed_aligned returns
base + off and the byte columns are the ones written here. Each case names the
branch of buff_simd.rs:317-2861 it reaches.
"""
import numpy as np

from buff_amd import datagen

SCALE = 1000
DLEN = 11
BASE = 1 << 30


def from_offsets(off):
    off = np.asarray(off, dtype=np.int64)
    return (BASE + off).astype(np.float64) / 2.0 ** DLEN


def bytes_col(rng, n, major, frac, lo=0, hi=256, avoid=None):
    """A byte column: `major` except for a `frac` share of rows drawn from [lo, hi) minus `avoid`."""
    col = np.full(n, major, dtype=np.int64)
    out = rng.random(n) < frac
    k = int(out.sum())
    v = rng.integers(lo, hi, k)
    bad = v == (major if avoid is None else avoid)
    v[bad] = (v[bad] + 1) % 256
    col[out] = v
    return col


def case_config3(n=60000, seed=4):
    """95% N(100, 0.5) and 5% +-10^U[6,9], 3 decimals: three sparse columns,
    two dense planes and a 2-bit tail (F = 42)."""
    return datagen.outlier_heavy(n, 3, seed)


def case_major_zero(n=20000, seed=11):
    """F = 24; column 0 majority 0x40, column 1 majority 0: a range target whose
    column-1 byte is 0 empties the candidates (buff_simd.rs:971-984)."""
    rng = np.random.default_rng(seed)
    c0 = bytes_col(rng, n, 0x40, 0.08)
    c1 = bytes_col(rng, n, 0x00, 0.06)
    c2 = rng.integers(0, 256, n)
    off = (c0 << 16) | (c1 << 8) | c2
    off[5] = 0
    off[7] = (1 << 24) - 1
    return from_offsets(off)


def case_max_quirks(n=20000, seed=12, below=False):
    """F = 24; the rows with the top column-0 byte (0xF0, outliers) have column-1
    byte equal to its majority (not ORed into the max, buff_simd.rs:2534-2538), or,
    with below=True, a column-1 outlier byte under the majority (the majority byte
    is chosen and the candidates empty)."""
    rng = np.random.default_rng(seed)
    c0 = bytes_col(rng, n, 0x40, 0.05, 0, 0xE0)
    c1 = bytes_col(rng, n, 0x30, 0.05)
    c2 = rng.integers(0, 256, n)
    top = rng.choice(np.arange(3000, n), 7, replace=False)
    c0[top] = 0xF0
    c1[top] = 0x05 if below else 0x30
    off = (c0 << 16) | (c1 << 8) | c2
    off[3] = 0
    return from_offsets(off)


def case_max_in_majority(n=30000, seed=18):
    """F = 24; every column-0 outlier byte is under the majority byte, so the
    majority rows hold the max (the staged max, not the row list)."""
    rng = np.random.default_rng(seed)
    c0 = bytes_col(rng, n, 0x80, 0.05, 0, 0x80)
    c1 = bytes_col(rng, n, 0x30, 0.05)
    c2 = rng.integers(0, 256, n)
    return from_offsets((c0 << 16) | (c1 << 8) | c2)


def case_max_many_ties(n=180000, seed=19):
    """F = 26; about 10,800 rows share the top column-0 outlier byte, more than
    the max's row list holds (the staged max runs instead)."""
    rng = np.random.default_rng(seed)
    c0 = np.where(rng.random(n) < 0.06, 0xF0, 0x40)
    c1 = bytes_col(rng, n, 0x30, 0.05)
    c2 = rng.integers(0, 1024, n)
    return from_offsets((c0 << 18) | (c1 << 10) | c2)


def case_all_sparse(n=12000, seed=13):
    """F = 16 with both byte columns sparse: decode returns nothing
    (buff_simd.rs:788-845), the sum keeps its base term."""
    rng = np.random.default_rng(seed)
    c0 = bytes_col(rng, n, 0x61, 0.04)
    c1 = bytes_col(rng, n, 0x17, 0.04)
    off = (c0 << 8) | c1
    off[0] = 0
    off[1] = 0xFFFF
    return from_offsets(off)


def case_no_outliers(n=9000, seed=14):
    """F = 24 with a column-1 byte of 0 on every row: that sparse column has no
    outliers and the decoder's unwrap panics (buff_simd.rs:683)."""
    rng = np.random.default_rng(seed)
    c0 = bytes_col(rng, n, 0x22, 0.05)
    c2 = rng.integers(0, 256, n)
    off = (c0 << 16) | c2
    off[2] = 0
    off[4] = (0xFF << 16) | 0xFF
    return from_offsets(off)


def case_no_majority(n=50000, seed=15):
    """A random walk: no top-byte majority, the stream is byte_fixed's behind a
    zero column count."""
    return datagen.random_walk(n, 3, seed)


def case_narrow_majority(n=3000, seed=16):
    """F < 8 with a majority value: the reference's `remain -= 8` underflows."""
    rng = np.random.default_rng(seed)
    off = np.where(rng.random(n) < 0.9, 5, rng.integers(0, 100, n))
    off[0] = 0
    off[1] = 99
    return from_offsets(off)


def case_narrow_plain(n=3000, seed=17):
    """F < 8 without a majority: written like byte_fixed (tail bits over the count
    byte), which the reference's readers cannot read back."""
    rng = np.random.default_rng(seed)
    off = rng.integers(0, 100, n)
    off[0] = 0
    off[1] = 99
    return from_offsets(off)


CASES = {
    "config3": (case_config3, SCALE),
    "config3_ragged": (lambda: datagen.outlier_heavy(100003, 3, 7), SCALE),
    "config3_min": (lambda: datagen.outlier_heavy(2000, 3, 8), SCALE),
    "major_zero": (case_major_zero, SCALE),
    "max_quirk_equal": (case_max_quirks, SCALE),
    "max_quirk_below": (lambda: case_max_quirks(below=True), SCALE),
    "all_sparse": (case_all_sparse, SCALE),
    "no_outliers": (case_no_outliers, SCALE),
    "no_majority": (case_no_majority, SCALE),
    "max_in_majority": (case_max_in_majority, SCALE),
    "max_many_ties": (case_max_many_ties, SCALE),
}


def predicates(x, seed=0):
    """Predicates drawn from the column plus the edges the readers special-case."""
    rng = np.random.default_rng(seed)
    xs = np.sort(x)
    picks = [float(x[i]) for i in rng.integers(0, len(x), 6)]
    qs = [float(np.quantile(x, q, method="lower")) for q in (0.001, 0.1, 0.5, 0.9, 0.999)]
    return picks + qs + [float(xs[0]), float(xs[-1]), float(xs[0]) - 1.0, float(xs[-1]) * 2 + 1.0, 100.0]
