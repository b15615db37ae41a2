"""refcsv — CPU restatement of the reference's CSV reader (test infrastructure).

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the device CSV
ingest; nothing in buff_amd/ imports it.

construct_file_iterator_skip_newline::<f64>(file, skip_val, delim)
(database/src/client.rs:163-182):
  BufReader::lines()   split on b"\\n"; a line's trailing "\\r" before the "\\n"
                       is removed (not a final "\\r" without "\\n"); a last line
                       without "\\n" still counts; invalid UTF-8 lines are
                       dropped (filter_map(Result::ok));
  line.split(delim)    every item, empty ones included;
  .skip(skip_val)      per line;
  item.parse::<f64>()  kept when Ok.

The parse is Rust's `f64::from_str` (Rust standard library, core::num::dec2flt,
1.55 and later: not part of the reference tree), restated from its documented
grammar — a string is accepted when, lowercased, it matches

    Float  ::= Sign? ( 'inf' | 'infinity' | 'nan' | Number )
    Number ::= ( Digit+ | Digit+ '.' Digit* | Digit* '.' Digit+ ) Exp?
    Exp    ::= 'e' Sign? Digit+
    Sign   ::= [+-]

— and its documented result: the nearest f64 (ties to even), infinity past the
range, signed zero below it; NaN for 'nan' (sign applied). Here the nearest f64
is computed from the exact rational value (fractions.Fraction and Python's
correctly rounded integer true division), not by CPython's float(str).
"""
import math
import re
import struct
from fractions import Fraction

_GRAMMAR = re.compile(r"[+-]?(?:inf|infinity|nan|(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:e[+-]?[0-9]+)?)\Z")
_MAX = Fraction(2) ** 1024 - Fraction(2) ** 970          # halfway between DBL_MAX and 2^1024
_MIN_SUB = Fraction(1, 2 ** 1074)


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def nearest_f64(r):
    """The f64 nearest to the non-negative rational r, ties to even."""
    if r == 0:
        return 0.0
    if r >= _MAX:
        return math.inf
    if r < 2 * _MIN_SUB:   # a multiple of 2^-1074 nearest to r, ties to even
        m = r / _MIN_SUB
        k = math.floor(m)
        frac = m - k
        if frac > Fraction(1, 2) or (frac == Fraction(1, 2) and k % 2 == 1):
            k += 1
        return _from_bits(k)
    # integer true division is correctly rounded
    return r.numerator / r.denominator


def parse_f64(item):
    """Rust's str::parse::<f64>() on bytes: the value, or None for Err."""
    try:
        s = item.decode("ascii")
    except UnicodeDecodeError:
        return None
    low = s.lower()
    if not _GRAMMAR.match(low):
        return None
    neg = low.startswith("-")
    body = low[1:] if low[:1] in "+-" else low
    if body in ("inf", "infinity"):
        v = math.inf
    elif body == "nan":
        return _from_bits(0xFFF8000000000000 if neg else 0x7FF8000000000000)
    else:
        mant, _, exp = body.partition("e")
        ip, _, fp = mant.partition(".")
        digits = (ip + fp) or "0"
        e = (int(exp) if exp else 0) - len(fp)
        if int(digits) == 0 or e < -400 - len(digits):
            v = 0.0 if int(digits) == 0 else nearest_f64(Fraction(0))
        elif e > 400:
            v = math.inf
        else:
            v = nearest_f64(Fraction(int(digits)) * Fraction(10) ** e)
    return -v if neg else v


def lines(text):
    """BufReader::lines() over bytes: the lines, without "\\n" / "\\r\\n"; None for a
    line that is not UTF-8 (dropped by the reader)."""
    out = []
    pos = 0
    n = len(text)
    while pos < n:
        j = text.find(b"\n", pos)
        if j < 0:
            line, pos = text[pos:], n
        else:
            line, pos = text[pos:j], j + 1
            if line.endswith(b"\r"):
                line = line[:-1]
        try:
            line.decode("utf-8")
        except UnicodeDecodeError:
            continue
        out.append(line)
    return out


def construct_file_iterator_skip_newline(text, skip_val=0, delim=b","):
    """The values the reference's iterator yields for a file holding `text`."""
    vals = []
    for line in lines(text):
        for item in line.split(delim)[skip_val:]:
            v = parse_f64(item)
            if v is not None:
                vals.append(v)
    return vals
