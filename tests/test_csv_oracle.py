"""Pins the CSV reader oracle (oracle/refcsv.py: client.rs:163-182 and Rust's
f64::from_str) against: the grammar Rust documents for f64::from_str (accepted
and rejected strings), CPython's float() (an independent correctly rounded
decimal parser) on grammar-valid items, values fixed by IEEE 754 (ties to even
at 2^53, the largest finite value and the overflow threshold, the smallest
subnormal and half of it, signed zero), BufReader::lines / str::split / skip
semantics on crafted texts, and the reference's own data file.
"""
import gzip
import math
import os
import struct

import numpy as np
import pytest

from oracle import refcsv as C

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "randomwalkdatasample1k-1k.csv.gz")


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


ACCEPT = ["1", "1.", ".5", "+1", "-0", "1e5", "1E5", "1e+5", "1.5E-3", "inf", "INF", "-Infinity", "NaN", "+nan",
          "00012", "0.000", "1.e3", "+.5e-2", "123456789012345678901234567890", "0e999999999999"]
REJECT = ["", ".", "e5", "1e", "1e+", "+", "-", " 1", "1 ", "1_0", "0x1F", "1.2.3", "--1", "infinit", "nana",
          "1e5.0", "1,0", "١", "+-1", "1e--5", "inf1", ".e5", "nan()"]


def test_grammar():
    for s in ACCEPT:
        assert C.parse_f64(s.encode("utf-8")) is not None, s
    for s in REJECT:
        assert C.parse_f64(s.encode("utf-8")) is None, s


def test_values_match_an_independent_parser():
    rng = np.random.default_rng(3)
    for _ in range(20000):
        nd = int(rng.integers(1, 26))
        digits = "".join(str(int(d)) for d in rng.integers(0, 10, nd))
        point = int(rng.integers(0, nd + 1))
        body = digits[:point] + "." + digits[point:] if rng.random() < 0.7 else digits
        if rng.random() < 0.6:
            body += "e" + str(int(rng.integers(-340, 320)))
        s = ("-" if rng.random() < 0.3 else "") + body
        got = C.parse_f64(s.encode())
        assert got is not None, s
        assert _bits(got) == _bits(float(s)), s


def test_values_fixed_by_ieee754():
    cases = {
        "9007199254740993": 9007199254740992.0,        # halfway above 2^53: ties to even
        "9007199254740995": 9007199254740996.0,
        "1.7976931348623157e308": 1.7976931348623157e308,
        "1.7976931348623158e308": 1.7976931348623157e308,
        "1.7976931348623159e308": math.inf,
        "4.9406564584124654e-324": 5e-324,
        "2.4703282292062327e-324": 0.0,                # below half the smallest subnormal
        "2.4703282292062328e-324": 5e-324,
        "2.2250738585072014e-308": 2.2250738585072014e-308,
        "1e400": math.inf,
        "1e-400": 0.0,
        "0.1": 0.1,
    }
    for s, want in cases.items():
        assert _bits(C.parse_f64(s.encode())) == _bits(want), s
    assert _bits(C.parse_f64(b"-0")) == 1 << 63
    assert _bits(C.parse_f64(b"-nan")) == 0xFFF8000000000000
    assert _bits(C.parse_f64(b"NaN")) == 0x7FF8000000000000


def test_lines_split_skip():
    f = C.construct_file_iterator_skip_newline
    assert f(b"1,2\r\n3\n\n4") == [1.0, 2.0, 3.0, 4.0]
    assert f(b"1\r") == []                                  # a final CR without LF stays: not a number
    assert f(b"1\r,2\n") == [2.0]                           # CR inside a line stays
    assert f(b"a,1,2\nb,3", skip_val=1) == [1.0, 2.0, 3.0]
    assert f(b"1,2,3\n4,5", skip_val=2) == [3.0]
    assert f(b"1,,2,\n", delim=b",") == [1.0, 2.0]
    assert f(b"1;2\n3", delim=b";") == [1.0, 2.0, 3.0]
    assert f(b"1,2\n\xff,3\n4") == [1.0, 2.0, 4.0]          # a line that is not UTF-8 is dropped
    assert f(b"1,\xc3\xa9,2") == [1.0, 2.0]                 # valid UTF-8: only the item fails
    assert f(b"") == [] and f(b"\n") == []


def test_reference_data_file():
    text = gzip.open(GOLDEN).read()
    vals = C.construct_file_iterator_skip_newline(text)
    assert len(vals) == 500 * 1000
    assert vals[:3] == [0.0, 0.53767, 2.3716]
    assert vals[-1] == -47.817
    items = [t for line in text.split(b"\n") for t in line.split(b",")]
    assert all(_bits(v) == _bits(float(t)) for v, t in zip(vals[::997], items[::997]))
