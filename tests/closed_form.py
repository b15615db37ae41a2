"""Closed-form restatement of the BUFF stream and query semantics in numpy.

Written from the format and the arithmetic the reference computes, not from
its control flow, so it can pin the line-by-line oracle (oracle/refcpu.cpp):
  * fixed(x) = trunc(x * 2^dlen) for finite |x| < 2^(63-dlen)
    (prec_double.rs:190-205 is a bit-level implementation of exactly this),
  * the `if/else if` running min/max (split_double.rs:321-330) is
    max{ f_i : f_i >= min(f_0..f_{i-1}) } with min of the empty prefix = i64::MAX,
  * the stream is a 20-byte header, floor(F/8) MSB-first byte planes and an
    LSB-first r-bit tail (split_double.rs:349-414),
  * a filter compares the stored F-bit offsets against the target taken
    modulo 2^F (the per-plane `as u8` truncation makes the digit-wise
    comparison equal to that), with the reference's early outs.
"""
import math

import numpy as np

PRECISION_MAP = {1: 5, 2: 8, 3: 11, 4: 15, 5: 18, 6: 21, 7: 25, 8: 28, 9: 31, 10: 35, 11: 38, 12: 50, 13: 10,
                 14: 10, 15: 10}
I64_MAX = (1 << 63) - 1
I64_MIN = -(1 << 63)


def dlen_of_scale(scale):
    p = int(round(math.log10(scale)))
    assert 10 ** p == scale
    return p, PRECISION_MAP[p]


def fixed(x, dlen):
    x = np.asarray(x, dtype=np.float64)
    assert np.all(np.abs(x) < 2.0 ** (62 - dlen))
    return np.trunc(np.ldexp(x, dlen)).astype(np.int64)


def quirk_minmax(f):
    """Running min / `else if` max of split_double.rs:321-330, vectorized."""
    f = np.asarray(f, dtype=np.int64)
    if len(f) == 0:
        return I64_MAX, I64_MIN
    excl = np.empty(len(f), dtype=np.int64)
    excl[0] = I64_MAX
    if len(f) > 1:
        excl[1:] = np.minimum.accumulate(f)[:-1]
    cand = f >= excl
    mx = int(f[cand].max()) if cand.any() else I64_MIN
    return int(f.min()), mx


def width(mn, mx):
    delta = (mx - mn) & ((1 << 64) - 1)
    if delta >= 1 << 63:
        delta -= 1 << 64
    if delta == 0:
        return 0
    if delta < 0:
        return 0
    return int(math.ceil(math.log2(float(delta))))


def encode(x, scale, flip=False):
    """Independent encoder: returns the reference byte stream (uint8)."""
    _, dlen = dlen_of_scale(scale)
    f = fixed(x, dlen)
    n = len(f)
    mn, mx = quirk_minmax(f)
    F = width(mn, mx)
    if F == 0 and n >= 1:
        raise ValueError("reference panics (F == 0)")
    off = (f.astype(np.uint64) - np.uint64(mn & ((1 << 64) - 1))) if n else np.zeros(0, np.uint64)
    hdr = np.zeros(20, dtype=np.uint8)
    hdr[0:8] = np.frombuffer(np.int64(mn).tobytes(), dtype=np.uint8)
    hdr[8:12] = np.frombuffer(np.uint32(n).tobytes(), dtype=np.uint8)
    hdr[12:16] = np.frombuffer(np.uint32((F - dlen) & 0xFFFFFFFF).tobytes(), dtype=np.uint8)
    hdr[16:20] = np.frombuffer(np.uint32(dlen).tobytes(), dtype=np.uint8)
    parts = [hdr]
    P, r = (F // 8, F % 8) if F >= 8 else (0, F)
    for p in range(P):
        b = ((off >> np.uint64(F - 8 * (p + 1))) & np.uint64(0xFF)).astype(np.uint8)
        parts.append(b ^ np.uint8(0x80) if flip else b)
    if r:
        bits = ((off[:, None] >> np.arange(r, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.uint8)
        parts.append(np.packbits(bits.reshape(-1), bitorder="little"))
    return np.concatenate(parts)


def parse(stream, flip=False):
    """Header fields and the stored F-bit offsets of a stream."""
    s = np.asarray(stream, dtype=np.uint8)
    base = int(np.frombuffer(s[0:8].tobytes(), dtype=np.int64)[0])
    n = int(np.frombuffer(s[8:12].tobytes(), dtype=np.uint32)[0])
    ilen = int(np.frombuffer(s[12:16].tobytes(), dtype=np.uint32)[0])
    dlen = int(np.frombuffer(s[16:20].tobytes(), dtype=np.uint32)[0])
    F = (ilen + dlen) & 0xFFFFFFFF
    P, r = (F // 8, F % 8) if F >= 8 else (0, F)
    off = np.zeros(n, dtype=np.uint64)
    pos = 20
    planes = []
    for p in range(P):
        b = s[pos:pos + n]
        if flip:
            b = b ^ np.uint8(0x80)
        planes.append(b)
        off = (off << np.uint64(8)) | b.astype(np.uint64)
        pos += n
    tail = np.zeros(n, dtype=np.uint64)
    if r:
        nb = (n * r + 7) // 8
        bits = np.unpackbits(s[pos:pos + nb], bitorder="little")[: n * r].reshape(n, r).astype(np.uint64)
        tail = (bits << np.arange(r, dtype=np.uint64)[None, :]).sum(axis=1).astype(np.uint64)
        off = (off << np.uint64(r)) | tail
    return dict(base=base, n=n, ilen=ilen, dlen=dlen, F=F, P=P, r=r, off=off, planes=planes, tail=tail)


def decode(stream, flip=False):
    h = parse(stream, flip)
    v = (np.uint64(h["base"] & ((1 << 64) - 1)) + h["off"]).view(np.int64)
    return v.astype(np.float64) / 2.0 ** h["dlen"]


def precision_exp(scale):
    p = math.log10(scale)
    prec = int(p)
    b = float("0." + "0" * prec + "49") if prec > 0 else 0.49
    return math.frexp(b)[1] - 1


def fixed_pred(pred, scale, dlen):
    pe = precision_exp(scale)
    e = math.frexp(abs(pred))[1] - 1 if pred != 0 else -1023
    if e < pe:
        return 0
    return int(math.trunc(math.ldexp(pred, dlen)))


def filter_bits(stream, scale, op, pred, semantic=False):
    """Boolean row mask the reference reports (or the intended one)."""
    h = parse(stream)
    n, F, base = h["n"], h["F"], h["base"]
    fp = fixed_pred(pred, scale, h["dlen"])
    if fp < base:
        return np.ones(n, bool) if op == "range" else np.zeros(n, bool)
    t = fp - base
    off = h["off"]
    if F < 8:
        if t >= 1 << 64:
            return np.zeros(n, bool)
        return (off > np.uint64(t)) if op == "range" else (off == np.uint64(t))
    if semantic and t >= (1 << F):
        return np.zeros(n, bool)
    tm = np.uint64(t % (1 << F))
    if op == "range":
        return off > tm
    if F % 8 == 0 and not semantic:
        return np.zeros(n, bool)
    return off == tm


def to_words(mask):
    """bool rows -> BitVec<u32> words."""
    n = len(mask)
    b = np.packbits(np.asarray(mask, dtype=np.uint8), bitorder="little")
    pad = (-len(b)) % 4
    b = np.concatenate([b, np.zeros(pad, np.uint8)])
    return b.view(np.uint32)[: (n + 31) // 32]


def exact_sum(stream):
    """Exact sum of the decoded values as a Python Fraction-free float pair."""
    from fractions import Fraction

    h = parse(stream)
    tot = h["n"] * h["base"] + int(h["off"].astype(object).sum()) if h["n"] else 0
    return Fraction(tot, 1 << h["dlen"]), Fraction(h["n"] * h["base"], 1 << h["dlen"])


def plane_sums(stream):
    h = parse(stream)
    out = [int(p.astype(np.uint64).sum()) for p in h["planes"]]
    if h["r"]:
        out.append(int(h["tail"].sum()))
    return out


def precision_R(h, precision):
    ilen, dlen = h["ilen"], h["dlen"]
    if precision == 0:
        R = ilen
        if ilen % 8:
            R = (ilen // 8 + 1) * 8
        return R & 0xFFFFFFFF
    bits = PRECISION_MAP[precision]
    if not dlen >= bits:
        raise ValueError("reference asserts dlen >= bits_needed")
    R = (ilen + bits) & 0xFFFFFFFF
    fixed_byte = ((dlen + ilen) & 0xFFFFFFFF) // 8
    bn, eb = R // 8, R % 8
    if bn < fixed_byte and eb > 0:
        R = (bn + 1) * 8
    return R


def decode_semantic_precision(stream, precision):
    """Values truncated to the top R stored bits, re-scaled."""
    h = parse(stream)
    R = min(precision_R(h, precision), h["F"])
    sh = np.uint64(h["F"] - R)
    off = (h["off"] >> sh) << sh
    v = (np.uint64(h["base"] & ((1 << 64) - 1)) + off).view(np.int64)
    return v.astype(np.float64) / 2.0 ** h["dlen"]
