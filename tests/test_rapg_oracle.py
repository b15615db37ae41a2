"""The RAPG residue format's oracle (oracle/refrapg.inc) against pins that do
not share its code.

* CRoaring's portable serialization of a run-optimized bitmap (croaring 0.5,
  absent from the reference tree): byte strings written out by hand from the
  Roaring format spec for array, run, bitset containers, the run/array and
  run/bitset size thresholds and the offset table (present without runs or
  with four or more containers), and random round trips through an independent
  Python reader (tests/closed_form.py).
* byte_residue_encode (split_double.rs:624-954): the stream parsed field by
  field in Python — header from the true minimum and maximum, the outlier
  columns from the leading-zero histogram of the first 2000 offsets
  (OUTLIER_R = 0.1), each column's bitmap rows and bytes = the rows whose byte
  is nonzero, one zero byte, the remaining planes and an LSB-first tail — must
  rebuild every row's offset; the reference's panics (fewer than 2000 rows with
  ilen > 15, the bit histogram read past its 32 slots) are reproduced.
* byte_residue_decode[_with_precision] / byte_residue_sum[_with_precision]
  (1353-1893, 2210-2629): a read model in Python (R bits, outlier columns ORed at
  R - 8(j+1), whole planes, then q-bit fields from the next byte; the sum
  leaves the base term out when the plane loop ends on a byte) must give the
  oracle's values bit for bit and its per-column u64 sums exactly.
"""
import numpy as np
import pytest

from buff_amd import datagen as dg
from tests import closed_form as cf

H = bytes.fromhex


@pytest.fixture(scope="module")
def R():
    from oracle import refcpu

    refcpu.build()
    return refcpu


ROARING_GOLDEN = [
    ([], "3a300000" "00000000"),
    ([5], "3a300000" "01000000" "0000" "0000" "10000000" "0500"),
    # 3 consecutive: a run (2 + 4 < 2*3 + 2); one container, so no offsets
    ([1, 2, 3], "3b300000" "01" "0000" "0200" "0100" "0100" "0200"),
    # 2 runs for 4 values: 2 + 8 == 2*4 + 2, so the array stays
    ([0, 1, 3, 4], "3a300000" "01000000" "0000" "0300" "10000000" "0000" "0100" "0300" "0400"),
    # a run and an array in two containers (keys 0 and 1)
    (list(range(10)) + [65536], "3b300100" "01" "0000" "0900" "0100" "0000" "0100" "0000" "0900" "0000"),
    # four containers with a run: the offset table is written
    ([0, 1, 2, 65536, 131072, 196608],
     "3b300300" "01" "0000" "0200" "0100" "0000" "0200" "0000" "0300" "0000"
     "25000000" "2b000000" "2d000000" "2f000000" "0100" "0000" "0200" "0000" "0000" "0000"),
    # 5000 consecutive: a bitset (card > 4096) that a run replaces (6 < 8192)
    (list(range(5000)), "3b300000" "01" "0000" "8713" "0100" "0000" "8713"),
]


@pytest.mark.parametrize("i", range(len(ROARING_GOLDEN)))
def test_roaring_golden(R, i):
    vals, hx = ROARING_GOLDEN[i]
    assert R.roaring_serialize(vals) == H(hx)
    assert cf.roaring_values(H(hx)) == vals
    assert list(R.roaring_deserialize(H(hx))) == vals


def test_roaring_bitset_stays(R):
    vals = list(range(0, 8193, 2))   # 4097 values, 4097 runs: 2 + 4 * 4097 >= 8192
    b = R.roaring_serialize(vals)
    words = np.zeros(1024, dtype="<u8")
    words[:128] = 0x5555555555555555
    words[128] = 1
    assert b == H("3a300000" "01000000" "0000" "0010" "10000000") + words.tobytes()
    assert cf.roaring_values(b) == vals


def test_roaring_random_round_trips(R):
    rng = np.random.default_rng(0)
    for _ in range(24):
        kind = rng.integers(0, 4)
        if kind == 0:
            v = np.unique(rng.integers(0, 1 << 20, int(rng.integers(0, 3000))))
        elif kind == 1:   # dense: bitsets
            v = np.unique(rng.integers(0, 1 << 17, int(rng.integers(5000, 12000))))
        elif kind == 2:   # runs
            starts = np.sort(rng.integers(0, 1 << 22, 40))
            v = np.unique(np.concatenate([np.arange(s, s + int(rng.integers(1, 1000))) for s in starts]))
        else:
            v = np.unique(np.concatenate([rng.integers(0, 70000, 100), np.arange(200000, 210000)]))
        b = R.roaring_serialize(v)
        assert cf.roaring_values(b) == [int(a) for a in v]
        assert np.array_equal(R.roaring_deserialize(b), v)


def _cases():
    return {
        "spikes_5pct": (dg.positive_spikes(60_000, 3, 1), 1000),
        "spikes_1pct": (dg.positive_spikes(100_003, 3, 2, 0.01), 1000),
        "spikes_dense": (dg.positive_spikes(5_000, 3, 3, 0.2), 1000),
        "spikes_p2": (dg.positive_spikes(40_000, 2, 4), 100),
        "spikes_p5": (dg.positive_spikes(30_000, 5, 5), 10 ** 5),
        "outlier_heavy": (dg.outlier_heavy(50_000, 3, 6), 1000),   # no outlier columns (negative outliers)
        "walk": (dg.random_walk(50_000, 3, 7), 1000),              # ilen <= 15: the byte_fixed layout
        "walk_p1": (dg.random_walk(9_000, 1, 8), 10),
        "constant": (np.full(3000, 2.5), 1000),                    # delta = 0: ilen wraps, 21 bytes
        "first_row_max": (np.concatenate([[2.0e7], dg.positive_spikes(20_000, 3, 10)]), 1000),
        "spikes_ragged_2001": (dg.positive_spikes(2001, 3, 11), 1000),   # one row past the 2000 sampled
        "wide_three_cols": (dg.wide_spikes(70_001, 3, 12), 1000),       # three outlier columns, a 1-bit tail
    }


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_encode_layout(R, name):
    x, scale = CASES[name]
    s, cols = R.residue_encode(x, scale)
    h = cf.rapg_parse(s)
    dlen = cf.dlen_of_scale(scale)[1]
    f = cf.fixed(x, dlen)
    mn, mx = int(f.min()), int(f.max())
    assert h["base"] == mn and h["n"] == len(x) and h["dlen"] == dlen
    F = cf.width(mn, mx)
    assert h["ilen"] == (F - dlen) & 0xFFFFFFFF and h["F"] == F & 0xFFFFFFFF
    off = (f - mn).astype(np.uint64)
    if name == "constant":
        assert len(s) == 21 and s[20] == 0
        return
    assert np.array_equal(h["off"], off)
    if h["ilen"] > 15:
        # the leading-zero histogram of the first 2000 offsets
        lz = np.array([64 - int(v).bit_length() for v in off[:2000]]) - (64 - F)
        cnt = np.bincount(lz[(lz >= 0) & (lz < 32)], minlength=32)
        want = 0
        for c in range(h["ilen"] // 8):
            if cnt[: 8 * (c + 1)].sum() < 200:
                want += 1
            else:
                break
        assert cols == want == h["S"]
        for j, c in enumerate(h["cols"]):
            byte = (off >> np.uint64(F - 8 * (j + 1))) & np.uint64(0xFF)
            rows = np.flatnonzero(byte)
            assert np.array_equal(c["rows"], rows) and np.array_equal(c["bytes"], byte[rows])
        if h["S"]:
            assert h["pad"] == 0
    else:
        assert cols == 0
    assert h["tail_at"] + (len(x) * h["r"] + 7) // 8 == len(s)   # nothing after the tail


@pytest.mark.parametrize("name", sorted(CASES))
def test_decode_and_sum_read_model(R, name):
    x, scale = CASES[name]
    s, _ = R.residue_encode(x, scale)
    for p in [-1, 0, 1, 2, 3, 4, 5, 6]:
        try:
            want = cf.rapg_decode(s, p)
        except (ValueError, KeyError, cf.Panic):
            with pytest.raises(R.RefPanic):
                R.residue_decode(s, p)
            with pytest.raises(R.RefPanic):
                R.residue_sum(s, p)
            continue
        got = R.residue_decode(s, p)
        assert len(got) == len(want), (name, p)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, p)
        sums, exact, _ = cf.rapg_sum(s, p)
        val, ps, cols = R.residue_sum(s, p)
        assert [int(v) for v in ps[: len(sums)]] == sums, (name, p)
        assert abs(val - float(exact)) <= 1e-12 * max(1.0, abs(float(exact))), (name, p, val, float(exact))
    if name != "constant":
        dlen = cf.dlen_of_scale(scale)[1]
        assert np.array_equal(R.residue_decode(s), cf.fixed(x, dlen) / 2.0 ** dlen)


def test_reference_panics(R):
    with pytest.raises(R.RefPanic):   # ilen > 15 samples fixed_vec[0..2000]
        R.residue_encode(dg.positive_spikes(1999, 3, 9), 1000)
    # ilen >= 40 with 2000 tiny offsets: the histogram read runs past bits_counter[31]
    x = np.zeros(5000)
    x[4000] = 2.0 ** 40
    with pytest.raises(R.RefPanic):
        R.residue_encode(x, 1000)
    # ilen = 34: zero_bytes * 8 = 32 slots, no panic; four outlier columns, the
    # last three without an outlier: the decode's outlier_iter.next().unwrap()
    # panics, the sum reads them as zero
    x[4000] = 1.5 * 2.0 ** 33
    s, cols = R.residue_encode(x, 1000)
    assert cols == 4
    with pytest.raises(R.RefPanic):
        R.residue_decode(s)
    sums, exact, _ = cf.rapg_sum(s)
    val, ps, _ = R.residue_sum(s)
    assert [int(v) for v in ps[: len(sums)]] == sums and val == float(exact)
