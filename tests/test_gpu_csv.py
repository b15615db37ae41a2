"""CSV ingest on the GPU through the C ABI (construct_file_iterator_skip_newline::<f64>,
client.rs:163-182) against the oracle, bit for bit: the reference's own data
file (and that file parsed on the device straight into the encoder, against the
oracle's stream of the oracle's values); mixed texts of valid and invalid items
with ',', ';', tab and newline-only layouts, CRLF lines, empty lines and items,
a final CR, skip_val 0-3; items of 1 to 25 significant digits and exponents
-340..320, and 20 to 300 digits (the exact big-integer path); values that
IEEE 754 fixes (ties to even, overflow, subnormals);
shortest-repr texts of 4M doubles parse back to the same bits; a file read by
buff_construct_file_iterator_skip_newline; the error cases (non-ASCII text,
a small output buffer).
"""
import gzip
import os
import struct
import tempfile

import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg
from oracle import refcsv as C

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "randomwalkdatasample1k-1k.csv.gz")


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _same(got, want):
    return len(got) == len(want) and np.array_equal(_bits(got), _bits(np.array(want, dtype=np.float64)))


def _rand_item(rng):
    r = rng.random()
    if r < 0.06:
        return rng.choice(["", ".", "e5", "1e", "1e+", "+", "-", "1_0", "0x1F", "1.2.3", "--1", "abc", "1e5.0",
                           "infinit", " 1", "1 "])
    if r < 0.09:
        return rng.choice(["inf", "-INF", "Infinity", "NaN", "+nan", "-0", "0", "0.000", "-0e10"])
    nd = int(rng.integers(1, 26))
    digits = "".join(str(int(d)) for d in rng.integers(0, 10, nd))
    lead = "0" * int(rng.integers(0, 3)) if rng.random() < 0.2 else ""
    point = int(rng.integers(0, nd + 1))
    body = lead + (digits[:point] + "." + digits[point:] if rng.random() < 0.7 else digits)
    if rng.random() < 0.5:
        body += rng.choice(["e", "E"]) + rng.choice(["", "+", "-"]) + str(int(rng.integers(0, 340)))
    return rng.choice(["", "-", "+"], p=[0.6, 0.3, 0.1]) + body


def _rand_text(rng, nlines, delim, crlf, final_nl):
    lines = []
    for _ in range(nlines):
        k = int(rng.integers(0, 12))
        lines.append(delim.join(_rand_item(rng) for _ in range(k)))
    nl = "\r\n" if crlf else "\n"
    t = nl.join(lines) + (nl if final_nl else "")
    return t.encode("ascii")


def test_reference_data_file(gpu_ctx):
    text = gzip.open(GOLDEN).read()
    want = C.construct_file_iterator_skip_newline(text)
    got = gpu_ctx.csv_parse(text)
    assert _same(got, want)


def test_ingest_into_the_encoder(gpu_ctx, oracle):
    # file bytes -> device parse -> device encode, against the oracle on the oracle's values
    text = gzip.open(GOLDEN).read()
    want = np.array(C.construct_file_iterator_skip_newline(text))
    tb = gpu_ctx.alloc(len(text))
    tb.upload(np.frombuffer(text, dtype=np.uint8))
    vb = gpu_ctx.alloc(len(want) * 8)
    try:
        n = gpu_ctx.csv_parse_device(tb.ptr, len(text), vb.ptr, len(want))
        assert n == len(want)
        for scale in (10 ** 4, 10 ** 5):
            ch = gpu_ctx.encode_device(vb.ptr, n, scale)
            try:
                assert np.array_equal(ch.serialize(), oracle.encode(want, scale))
            finally:
                ch.free()
    finally:
        tb.free()
        vb.free()


@pytest.mark.parametrize("seed,delim,crlf,final_nl,skip", [
    (1, ",", False, True, 0), (2, ",", True, False, 0), (3, ";", False, False, 1), (4, "\t", True, True, 3),
    (5, ",", False, True, 2), (6, "\n", False, True, 0), (7, ",", True, True, 0), (8, " ", False, False, 1)])
def test_mixed_texts(gpu_ctx, seed, delim, crlf, final_nl, skip):
    rng = np.random.default_rng(seed)
    text = _rand_text(rng, 3000, delim, crlf, final_nl)
    if seed == 7:
        text += b"\r"                                   # a final CR without LF
    want = C.construct_file_iterator_skip_newline(text, skip_val=skip, delim=delim.encode())
    got = gpu_ctx.csv_parse(text, delim=delim.encode(), skip_val=skip)
    assert _same(got, want), (seed, len(got), len(want))


def test_values_fixed_by_ieee754(gpu_ctx):
    items = ["9007199254740993", "9007199254740995", "1.7976931348623157e308", "1.7976931348623158e308",
             "1.7976931348623159e308", "4.9406564584124654e-324", "2.4703282292062327e-324",
             "2.4703282292062328e-324", "2.2250738585072014e-308", "2.2250738585072011e-308", "1e400", "1e-400",
             "0.1", "123456789012345678901234567890", "0.30000000000000004", "5e-324", "1e23", "8.98846567431158e307",
             "179769313486231580793728971405301e276", "-0", "0e999999999999", "1e-99999999999"]
    for it in items:
        want = C.construct_file_iterator_skip_newline(it.encode())
        assert _same(gpu_ctx.csv_parse(it.encode()), want), it
    text = ",".join(items).encode()
    assert _same(gpu_ctx.csv_parse(text), C.construct_file_iterator_skip_newline(text))


def test_long_items_take_the_exact_path(gpu_ctx):
    # 20 to 60 significant digits: items whose 19-digit bounds straddle a
    # rounding boundary are decided by the big-integer path
    rng = np.random.default_rng(12)
    toks = []
    for _ in range(60000):
        nd = int(rng.integers(20, 61))
        digits = "".join(str(int(d)) for d in rng.integers(0, 10, nd))
        point = int(rng.integers(0, nd + 1))
        body = digits[:point] + "." + digits[point:]
        if rng.random() < 0.6:
            body += "e" + str(int(rng.integers(-360, 330)))
        toks.append(("-" if rng.random() < 0.3 else "") + body)
    text = ",".join(toks).encode()
    assert _same(gpu_ctx.csv_parse(text), C.construct_file_iterator_skip_newline(text))
    # just above the tie between 2^53 and 2^53 + 2: the 19-digit bounds disagree,
    # so the exact path decides it, up to 300 significant digits
    at300 = b"9007199254740993." + b"0" * 283 + b"1"
    assert _same(gpu_ctx.csv_parse(at300), C.construct_file_iterator_skip_newline(at300))
    assert gpu_ctx.csv_parse(at300)[0] == 9007199254740994.0
    with pytest.raises(N.Unsupported):
        gpu_ctx.csv_parse(b"1," + b"9007199254740993." + b"0" * 290 + b"1" + b",2")


def test_shortest_repr_round_trip(gpu_ctx):
    rng = np.random.default_rng(9)
    n = 1 << 22
    x = dg.random_walk(n, 6, 3) * np.exp(rng.uniform(-30, 30, n))
    x[::1000] = rng.uniform(-1e300, 1e300, len(x[::1000]))
    text = ",".join(map(repr, x.tolist())).encode()
    got = gpu_ctx.csv_parse(text)
    assert np.array_equal(_bits(got), _bits(x))


def test_file_entry_point_and_errors(gpu_ctx):
    text = b"0,1.5,-2.25\r\n3e2,x,4\n"
    with tempfile.NamedTemporaryFile(suffix=".csv", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        assert _same(N.construct_file_iterator_skip_newline(path), [0.0, 1.5, -2.25, 300.0, 4.0])
        assert _same(N.construct_file_iterator_skip_newline(path, skip_val=1), [1.5, -2.25, 4.0])
    finally:
        os.unlink(path)
    with pytest.raises(N.BuffError):
        N.construct_file_iterator_skip_newline(path)      # removed: the reference's Err(())
    with pytest.raises(N.Unsupported):
        gpu_ctx.csv_parse(b"1,\xc3\xa9,2")
    with pytest.raises(N.BuffError):
        gpu_ctx.csv_parse(b"1,2", delim=b"\r")
    tb = gpu_ctx.alloc(64)
    vb = gpu_ctx.alloc(64)
    try:
        t = np.frombuffer(b"1,2,3,4,5", dtype=np.uint8)
        tb.upload(t)
        with pytest.raises(N.BuffError):
            gpu_ctx.csv_parse_device(tb.ptr, len(t), vb.ptr, 3)
        assert gpu_ctx.csv_parse_device(tb.ptr, len(t), vb.ptr, 5) == 5
        assert _same(vb.download(np.float64, 5), [1, 2, 3, 4, 5])
    finally:
        tb.free()
        vb.free()
