"""Pins the oracle's buff-major restatement (oracle/refcpu.cpp, buff_simd.rs:317-2861)
against the closed-form model in closed_form.py, and its snappy against the format.

The closed form is written from what each reader computes per row (including the
reference's sparse-column rules), not from its control flow; the oracle walks the
byte stream with BitPack cursors exactly like the Rust code. Agreement on every
case and predicate pins both.
"""
import numpy as np
import pytest

from tests import closed_form as cf
from tests import majority_cases as mc
from oracle import refcpu as R


@pytest.fixture(scope="module", params=sorted(mc.CASES))
def case(request):
    build, scale = mc.CASES[request.param]
    x = build()
    s = R.major_encode(x, scale)
    return request.param, x, scale, s, cf.major_parse(s)


def test_layout_matches_definition(case):
    name, x, scale, s, h = case
    lay = cf.major_layout(x, scale)
    assert h["base"] == lay["base"] and h["F"] == lay["F"] and h["dlen"] == lay["dlen"]
    assert h["S"] == lay["S"] and h["majors"] == lay["majors"], (h["majors"], lay["majors"])
    assert len(h["cols"]) == len(lay["cols"])
    for a, b in zip(h["cols"], lay["cols"]):
        assert np.array_equal(a, b)
    assert np.array_equal(h["tail"], lay["tail"])
    # every outlier byte differs from its column's majority byte
    for j in range(h["S"]):
        assert np.all(h["outliers"][j] != h["majors"][j])


def test_expected_shapes():
    """The crafted cases reach the branches their names promise."""
    want = {"config3": (42, 3), "major_zero": (24, 2), "max_quirk_equal": (24, 2), "all_sparse": (16, 2),
            "no_outliers": (24, 2), "no_majority": (None, 0)}
    for name, (F, S) in want.items():
        build, scale = mc.CASES[name]
        h = cf.major_parse(R.major_encode(build(), scale))
        assert h["S"] == S, name
        if F is not None:
            assert h["F"] == F, name
    h = cf.major_parse(R.major_encode(mc.case_no_outliers(), mc.SCALE))
    assert int(h["masks"][1].sum()) == 0


def test_decode(case):
    name, x, scale, s, h = case
    try:
        want = cf.major_decode(h)
    except ValueError:
        with pytest.raises(R.RefPanic):
            R.major_decode(s, scale)
        return
    got = R.major_decode(s, scale)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    if name == "all_sparse":
        assert len(got) == 0


def test_filters(case):
    name, x, scale, s, h = case
    for pred in mc.predicates(x):
        for op in ("range", "equal"):
            words, cnt = R.major_filter(s, scale, op, pred)
            m = cf.major_filter(h, scale, op, pred)
            assert np.array_equal(words, cf.to_words(m)), (name, op, pred)
            assert cnt == int(m.sum())


def test_sum(case):
    name, x, scale, s, h = case
    got, cs = R.major_sum(s, scale)
    want, sums = cf.major_sum(h)
    assert got == want or (np.isnan(got) and np.isnan(want))
    assert list(cs[: len(sums)]) == sums


def test_max(case):
    name, x, scale, s, h = case
    assert R.major_max(s, scale) == cf.major_max(h)


def test_quirks_are_visible():
    """The reference rules the model encodes change answers on these columns."""
    # sparse-column range rule: the target's column-1 byte is 0 = majority byte
    x = mc.case_major_zero()
    s = R.major_encode(x, mc.SCALE)
    h = cf.major_parse(s)
    off = (np.round(x * 2.0 ** mc.DLEN).astype(np.int64) - h["base"]).astype(np.uint64)
    row = int(np.flatnonzero((h["cols"][0] == 0x40) & (h["cols"][1] == 0))[0])
    pred = float(x[row]) - 2.0 ** -mc.DLEN * 50
    words, cnt = R.major_filter(s, mc.SCALE, "range", pred)
    t = int(off[row]) - 50
    assert cnt < int((off > np.uint64(t)).sum())
    # max: candidates all at the column-1 majority byte -> that byte is left out
    x = mc.case_max_quirks()
    s = R.major_encode(x, mc.SCALE)
    v, mo, _ = R.major_max(s, mc.SCALE)
    assert v < float(x.max())
    # decode: the second sparse column ORs its majority byte at twice its shift
    x = mc.case_config3()
    d = R.major_decode(R.major_encode(x, mc.SCALE), mc.SCALE)
    assert np.mean(d == x) < 0.01


@pytest.mark.parametrize("build", [mc.case_narrow_majority, lambda: mc.case_config3(1999)])
def test_encode_panics(build):
    with pytest.raises(R.RefPanic):
        R.major_encode(build(), mc.SCALE)


def test_narrow_plain_stream_is_byte_fixed():
    """F < 8 without a majority: the stream is byte_fixed_encode's (tail bits
    over the zero count byte), and the majority readers panic on it."""
    x = mc.case_narrow_plain()
    s = R.major_encode(x, mc.SCALE)
    assert np.array_equal(s, R.encode(x, mc.SCALE))
    with pytest.raises(R.RefPanic):
        R.major_decode(s, mc.SCALE)
    with pytest.raises(R.RefPanic):
        R.major_sum(s, mc.SCALE)


def test_no_majority_stream_is_byte_fixed_behind_count():
    x = mc.case_no_majority()
    s = R.major_encode(x, mc.SCALE)
    plain = R.encode(x, mc.SCALE)
    # plain max uses the `else if` max; this column's first row is not its maximum
    assert np.array_equal(s[:20], plain[:20]) and s[20] == 0 and np.array_equal(s[21:], plain[20:])
    assert np.array_equal(R.major_decode(s, mc.SCALE).view(np.int64), R.decode(plain, mc.SCALE).view(np.int64))


# ---- snappy -------------------------------------------------------------------

def _inputs():
    rng = np.random.default_rng(3)
    yield b""
    yield b"a"
    yield bytes(range(14))
    yield rng.integers(0, 256, 100000, dtype=np.uint8).tobytes()   # incompressible
    yield bytes(70000)                                              # one run over two fragments
    bits = (rng.random(8 * 40000) < 0.05).astype(np.uint8)
    yield np.packbits(bits, bitorder="little").tobytes()
    yield rng.integers(0, 4, 200000, dtype=np.uint8).tobytes()


def test_snappy_roundtrip_and_format():
    for d in _inputs():
        c = R.snappy_compress(d)
        assert R.snappy_decompress(c, len(d)) == d
        assert cf.snappy_decompress(c) == d


def test_snappy_literal_known_answers():
    # varint length, then one literal tag (len-1)<<2 for inputs below the 15-byte margin
    assert R.snappy_compress(b"") == b"\x00"
    assert R.snappy_compress(b"a") == b"\x01\x00a"
    assert R.snappy_compress(bytes(range(14))) == b"\x0e\x34" + bytes(range(14))
    # 100 zero bytes: literal of 1, then copies of offset 1 (64 + 35 bytes, 2-byte offsets)
    assert R.snappy_compress(bytes(100)) == b"\x64\x00\x00\xfe\x01\x00\x8a\x01\x00"


def test_snappy_against_pyarrow():
    """pyarrow ships Google snappy: its decoder reads our streams, ours reads its
    streams, and on input where every release emits the same bytes
    (incompressible data, runs) the compressed bytes agree too."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(5)
    for d in _inputs():
        if not d:
            continue
        c = R.snappy_compress(d)
        assert pa.decompress(c, decompressed_size=len(d), codec="snappy", asbytes=True) == d
        theirs = pa.compress(d, codec="snappy", asbytes=True)
        assert R.snappy_decompress(theirs, len(d)) == d
    for d in (rng.integers(0, 256, 300000, dtype=np.uint8).tobytes(), bytes(300000), bytes(100), b"x" * 5000):
        assert R.snappy_compress(d) == pa.compress(d, codec="snappy", asbytes=True)


def test_condition_forms(case):
    """buff_range_filter_majority_condition / buff_major_decode_condition
    (buff_simd.rs:1279-1526, 1714-1914): with sparse columns both give nothing
    (the reference logs "not implemented"); without, the stream is byte_fixed's
    behind a zero column-count byte, and they answer as the plain conditional
    forms on that stream, except that the decode places the tail bits correctly
    and an empty condition gives no values."""
    name, x, scale, s, h = case
    n = h["n"]
    plain = np.concatenate([s[:20], s[21:]]) if h["S"] == 0 else None
    rng = np.random.default_rng(21)
    for frac in (0.0, 0.01, 0.5, 1.0):
        cond = rng.random(n) < frac
        rows = np.flatnonzero(cond)
        for pred in mc.predicates(x)[:8]:
            words, cnt = R.filter_cond(s, scale, pred, rows, major=True)
            want = np.zeros(n, bool) if plain is None else cf.filter_cond_bits(plain, scale, "range", pred, cond)
            assert np.array_equal(words, cf.to_words(want)), (name, frac, pred)
            assert cnt == int(want.sum())
        got = R.decode_cond(s, scale, rows, major=True)
        want = np.zeros(0) if plain is None else cf.decode_cond(plain, cond, semantic=True)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, frac)


def test_max_range_majority(case):
    """buff_max_majority_range (buff_simd.rs:2863-3114): nothing with sparse
    columns; without, buff_max_range on the stream without its column-count byte."""
    name, x, scale, s, h = case
    n = h["n"]
    plain = np.concatenate([s[:20], s[21:]]) if h["S"] == 0 else None
    for a, b, w in ((0, n, n), (0, n - n % 100, 100), (5, 5 + 1000 * 7, 7), (3, n, 1000)):
        if plain is None:
            offs, vals, words = R.max_range(s, a, b, w, major=True)
            assert len(offs) == 0 and not words.any(), name
            continue
        try:
            want = R.max_range(plain, a, b, w)
        except R.RefPanic:
            with pytest.raises(R.RefPanic):
                R.max_range(s, a, b, w, major=True)
            continue
        got = R.max_range(s, a, b, w, major=True)
        for g, wv in zip(got, want):
            assert np.array_equal(g, wv), (name, a, b, w)
