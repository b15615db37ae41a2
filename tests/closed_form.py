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


def filter_cond_bits(stream, scale, op, pred, cond, semantic=False):
    """Rows of `cond` the conditional filter reports. buff_range_filter_condition
    (buff_simd.rs:4683-4921) returns no rows for a predicate below the base and
    compares against the target's low byte when F < 8; otherwise (and with
    semantic) it is the plain filter on the condition rows."""
    h = parse(stream)
    cond = np.asarray(cond, bool)
    if op == "range" and not semantic:
        fp = fixed_pred(pred, scale, h["dlen"])
        if fp < h["base"]:
            return np.zeros(h["n"], bool)
        if 0 < h["F"] < 8:
            return (h["off"] > np.uint64((fp - h["base"]) & 0xFF)) & cond
    return filter_bits(stream, scale, op, pred, semantic=semantic) & cond


def decode_cond(stream, cond, semantic=False):
    """Values of the `cond` rows in row order (buff_decode_condition,
    buff_simd.rs:1528-1712): offset o = planes * 2^r + tail, but the reference
    ORs the tail bits in at bit r when F >= 8 and r > 0: (planes | tail) * 2^r."""
    h = parse(stream)
    rows = np.flatnonzero(np.asarray(cond, bool))
    off = h["off"][rows].astype(np.uint64)
    F, r = h["F"], h["F"] % 8
    if not semantic and F >= 8 and r:
        rr = np.uint64(r)
        off = ((off >> rr) | (off & np.uint64((1 << r) - 1))) << rr
    return (np.int64(h["base"]) + off.astype(np.int64)).astype(np.float64) / float(1 << h["dlen"])


class Panic(Exception):
    """The reference panics on this input."""


def max_range(stream, s, e, window, semantic=False):
    """buff_max_range (buff_simd.rs:5096-5334) from its arithmetic: the max stored
    offset of each window of `window` rows of [s, e) and the rows holding it.
    The reference keeps only full windows, panics when F > 8 and the windows do
    not tile [s, e) exactly (or it is empty), and fills the argmax rows only in
    its tail stage (F < 8, or F > 8 with a tail); semantic: every window, every
    argmax row. Returns (max offsets, argmax row mask)."""
    h = parse(stream)
    off, F, n = h["off"], h["F"], h["n"]
    L = e - s
    nall = -(-L // window) if L > 0 else 0
    seg = off[s:e]
    wm = np.array([int(seg[k * window:(k + 1) * window].max()) for k in range(nall)], dtype=np.uint64)
    rows = np.arange(n)
    inr = (rows >= s) & (rows < e)
    widx = np.where(inr, (rows - s) // window, 0)
    arg = np.zeros(n, bool)
    if nall:
        arg[inr] = off[inr] == wm[widx[inr]]
    if semantic:
        return wm, arg
    if F > 8 and (L == 0 or L % window):
        raise Panic("assert_eq! / get_mut().unwrap()")
    if F == 8 or (F > 8 and F % 8 == 0):
        arg[:] = False
    return wm[: L // window], arg


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


def sum_precision(stream, precision):
    """What byte_residue_sum_with_precision reads and returns on a plain stream,
    from the stream's bytes: the R-bit read rule gives R // 8 whole n-byte
    sub-columns after the header, then (R % 8 = q > 0) n q-bit LSB-first fields
    taken from the next byte on. Each sub-column k weighs 2^(ilen - 8(k+1)), the
    fields 2^(ilen - R), and the n * base / 2^dlen term is left out when the read
    ends on a byte boundary (R >= 8, R % 8 == 0). Returns (sums, exact value as
    a Fraction); raises Panic when the reads run past the stream."""
    from fractions import Fraction

    s = np.asarray(stream, dtype=np.uint8)
    h = parse(stream)
    n = h["n"]
    R = precision_R(h, precision)
    ilen_s = h["F"] - h["dlen"]           # ilen read as i32 (it wraps below 0)
    nplanes, q = R // 8, R % 8
    if 20 + nplanes * n > len(s):
        raise Panic("whole sub-columns past the stream")
    sums = [int(s[20 + k * n: 20 + (k + 1) * n].astype(np.uint64).sum()) for k in range(nplanes)]
    val = sum(Fraction(v) * Fraction(2) ** (ilen_s - 8 * (k + 1)) for k, v in enumerate(sums))
    if q:
        start = 20 + nplanes * n
        nb = (n * q + 7) // 8
        if start + nb > len(s):
            raise Panic("fields past the stream")
        bits = np.unpackbits(s[start:start + nb], bitorder="little")[: n * q].reshape(n, q).astype(np.uint64)
        fs = int((bits << np.arange(q, dtype=np.uint64)[None, :]).sum())
        sums.append(fs)
        val += Fraction(fs) * Fraction(2) ** (ilen_s - R)
    if not (R >= 8 and q == 0):
        val += Fraction(n * h["base"], 1 << h["dlen"])
    return sums, val


def sum_semantic_precision(stream, precision):
    """Exact sum of the values truncated to their top min(R, F) stored bits."""
    from fractions import Fraction

    h = parse(stream)
    R = min(precision_R(h, precision), h["F"])
    sh = np.uint64(h["F"] - R)
    off = (h["off"] >> sh) << sh
    return Fraction(h["n"] * h["base"] + int(off.astype(object).sum()), 1 << h["dlen"])


def decode_semantic_precision(stream, precision):
    """Values truncated to the top R stored bits, re-scaled."""
    h = parse(stream)
    R = min(precision_R(h, precision), h["F"])
    sh = np.uint64(h["F"] - R)
    off = (h["off"] >> sh) << sh
    v = (np.uint64(h["base"] & ((1 << 64) - 1)) + off).view(np.int64)
    return v.astype(np.float64) / 2.0 ** h["dlen"]


# --- buff-major (buff_simd.rs:317-2861) -------------------------------------
# Stream: header | u8 ncols | per sparse column [u8 major][u32 bm_len]
# [snappy(u64 nbits | u32 words)][u32 n_out][n_out outlier bytes] | one zero byte
# (only when ncols > 0) | dense MSB-first planes | LSB-first tail.

SAMPLE, M_LIMIT = 2000, 1200


def snappy_decompress(buf):
    """Raw snappy decoder written from the format description."""
    buf = bytes(buf)
    pos, n, shift = 0, 0, 0
    while True:
        b = buf[pos]
        pos += 1
        n |= (b & 0x7F) << shift
        if not b & 0x80:
            break
        shift += 7
    out = bytearray()
    while pos < len(buf):
        tag = buf[pos]
        pos += 1
        kind = tag & 3
        if kind == 0:
            ln = (tag >> 2) + 1
            if ln > 60:
                nb = ln - 60
                ln = int.from_bytes(buf[pos:pos + nb], "little") + 1
                pos += nb
            out += buf[pos:pos + ln]
            pos += ln
            continue
        if kind == 1:
            ln, off = ((tag >> 2) & 7) + 4, ((tag >> 5) << 8) | buf[pos]
            pos += 1
        elif kind == 2:
            ln, off = (tag >> 2) + 1, int.from_bytes(buf[pos:pos + 2], "little")
            pos += 2
        else:
            ln, off = (tag >> 2) + 1, int.from_bytes(buf[pos:pos + 4], "little")
            pos += 4
        assert 0 < off <= len(out)
        if off >= ln:
            out += out[len(out) - off:len(out) - off + ln]
        else:
            for _ in range(ln):
                out.append(out[-off])
    assert len(out) == n
    return bytes(out)


def boyer_moore(values):
    major, freq = 0, 0
    for v in values:
        if freq == 0:
            major, freq = v, 1
        elif v == major:
            freq += 1
        else:
            freq -= 1
    return major, freq


def major_layout(x, scale):
    """What buff_encode_majority stores, from its definition: base/F from the
    plain min and max, the majority bytes from the Boyer-Moore vote over the
    first 2000 offsets (accepted while the counter exceeds 1200), then every
    byte column and the tail."""
    _, dlen = dlen_of_scale(scale)
    f = fixed(x, dlen)
    n = len(f)
    if n < SAMPLE:
        raise ValueError("reference panics (n < 2000)")
    mn, mx = int(f.min()), int(f.max())
    F = width(mn, mx)
    off = f.astype(np.uint64) - np.uint64(mn & ((1 << 64) - 1))
    head = [int(v) for v in off[:SAMPLE]]
    majors = []
    rsf = F - 8 if F >= 8 else 0
    m, fr = boyer_moore([(v >> rsf) & 0xFF for v in head])
    if fr > M_LIMIT:
        majors.append(m)
        while rsf > 7:
            rsf -= 8
            m, fr = boyer_moore([(v >> rsf) & 0xFF for v in head])
            if fr > M_LIMIT:
                majors.append(m)
            else:
                break
    S = len(majors)
    if S and F < 8:
        raise ValueError("reference does not terminate (F < 8 with a majority byte)")
    cols = [((off >> np.uint64(F - 8 * (j + 1))) & np.uint64(0xFF)).astype(np.uint8) for j in range((F // 8) if F >= 8 else 0)]
    rr = F - 8 * len(cols)
    tail = (off & np.uint64((1 << rr) - 1)) if rr else np.zeros(n, np.uint64)
    return dict(base=mn, n=n, F=F, dlen=dlen, ilen=(F - dlen) & 0xFFFFFFFF, S=S, majors=majors, cols=cols,
                r=rr, tail=tail)


def major_parse(stream):
    """Parse a buff-major stream into its byte columns (outliers scattered back)."""
    s = bytes(np.asarray(stream, dtype=np.uint8).tobytes())
    base = int.from_bytes(s[0:8], "little", signed=True)
    n = int.from_bytes(s[8:12], "little")
    ilen = int.from_bytes(s[12:16], "little")
    dlen = int.from_bytes(s[16:20], "little")
    F = (ilen + dlen) & 0xFFFFFFFF
    S = s[20]
    pos = 21
    majors, masks, cols, bm_lens, outliers = [], [], [], [], []
    for _ in range(S):
        major = s[pos]
        bm_len = int.from_bytes(s[pos + 1:pos + 5], "little")
        binary = snappy_decompress(s[pos + 5:pos + 5 + bm_len])
        pos += 5 + bm_len
        nout = int.from_bytes(s[pos:pos + 4], "little")
        ob = np.frombuffer(s[pos + 4:pos + 4 + nout], dtype=np.uint8)
        pos += 4 + nout
        nbits = int.from_bytes(binary[:8], "little")
        assert nbits == n and len(binary) == 8 + 4 * ((n + 31) // 32)
        mask = np.unpackbits(np.frombuffer(binary[8:], dtype=np.uint8), bitorder="little")[:n].astype(bool)
        assert int(mask.sum()) == nout
        col = np.full(n, major, dtype=np.uint8)
        col[mask] = ob
        majors.append(major)
        masks.append(mask)
        cols.append(col)
        bm_lens.append(bm_len)
        outliers.append(ob)
    if S:
        assert s[pos] == 0
        pos += 1
    remain = F - 8 * S
    P, r = remain // 8, remain % 8
    for _ in range(P):
        cols.append(np.frombuffer(s[pos:pos + n], dtype=np.uint8))
        pos += n
    tail = np.zeros(n, dtype=np.uint64)
    if r:
        nb = (n * r + 7) // 8
        bits = np.unpackbits(np.frombuffer(s[pos:pos + nb], dtype=np.uint8), bitorder="little")[: n * r]
        tail = (bits.reshape(n, r).astype(np.uint64) << np.arange(r, dtype=np.uint64)[None, :]).sum(axis=1)
        tail = tail.astype(np.uint64)
        pos += nb
    assert pos == len(s), (pos, len(s))
    return dict(base=base, n=n, ilen=ilen, dlen=dlen, F=F, S=S, majors=majors, masks=masks, cols=cols, r=r,
                tail=tail, bm_lens=bm_lens, outliers=outliers)


def _shift(h, j):
    return h["F"] - 8 * (j + 1)


def _as_f64(h, fixed_u64):
    v = (np.uint64(h["base"] & ((1 << 64) - 1)) + fixed_u64).view(np.int64)
    return v.astype(np.float64) / 2.0 ** h["dlen"]


def major_decode(h):
    """buff_decode_majority's output: sparse columns after the first OR the
    majority byte shifted twice, and only for rows before that column's last
    outlier; no output when the sparse columns end the stream; a sparse column
    without outliers is an unwrap panic."""
    n, S = h["n"], h["S"]
    M64 = (1 << 64) - 1
    fx = np.zeros(n, dtype=np.uint64)
    for j in range(S):
        sh = _shift(h, j)
        mask = h["masks"][j]
        if not mask.any():
            raise ValueError("reference panics (sparse column without outliers)")
        b = h["cols"][j].astype(np.uint64) << np.uint64(sh)
        if j == 0:
            fx = b
            continue
        last = int(np.flatnonzero(mask)[-1])
        twice = np.uint64(((h["majors"][j] << sh) << sh) & M64)
        part = np.where(mask, b, twice)
        part[last + 1:] = 0
        fx |= part
    remain = h["F"] - 8 * S
    if S and remain == 0:
        return np.zeros(0, dtype=np.float64)
    for k in range(S, len(h["cols"])):
        fx |= h["cols"][k].astype(np.uint64) << np.uint64(_shift(h, k))
    if h["r"]:
        fx |= h["tail"]
    return _as_f64(h, fx)


def major_filter(h, scale, op, pred):
    """Row mask of buff_range_filter_majority / buff_equal_filter_majority."""
    n, S = h["n"], h["S"]
    fp = fixed_pred(pred, scale, h["dlen"])
    if fp < h["base"]:
        return np.ones(n, bool) if op == "range" else np.zeros(n, bool)
    t = (fp - h["base"]) & ((1 << 64) - 1)
    res = np.zeros(n, bool)
    cand = np.ones(n, bool)
    for j, col in enumerate(h["cols"]):
        tb = (t >> _shift(h, j)) & 0xFF
        if op == "equal":
            cand &= col == tb
            continue
        if j >= S or j == 0:
            res |= cand & (col > tb)
            cand &= col == tb
            continue
        mask, major = h["masks"][j], h["majors"][j]
        if tb == 0 and major == 0:            # candidates are dropped
            res |= cand & mask
            cand[:] = False
        else:
            res |= mask & (col > tb)          # outliers: not masked by the candidates
            if major > tb:
                res |= cand & ~mask
            cand &= col == tb
    if h["r"]:
        tt = t & ((1 << h["r"]) - 1)
        hit = h["tail"] > np.uint64(tt) if op == "range" else h["tail"] == np.uint64(tt)
        res |= cand & hit
    return res


def major_sum(h):
    """buff_sum_majority: the per-column fold; the base term is dropped when a
    dense plane ends the stream."""
    ilen_i32 = h["ilen"] - (1 << 32) if h["ilen"] >= 1 << 31 else h["ilen"]
    total, processed = 0.0, 0
    sums = [int(c.astype(np.uint64).sum()) for c in h["cols"]]
    for k, sk in enumerate(sums):
        processed += 8
        total = total + float(sk) / 2.0 ** (processed - ilen_i32)
        if k >= h["S"] and k == len(sums) - 1 and h["r"] == 0:
            return total, sums
    if h["r"]:
        processed += h["r"]
        st = int(h["tail"].sum())
        sums.append(st)
        total = total + float(st) / 2.0 ** (processed - ilen_i32)
    base_term = float(h["n"]) * float(h["base"]) / 2.0 ** h["dlen"]
    return total + base_term, sums


def major_max(h):
    """buff_max_majority: the byte-wise progressive max with its sparse-column rules."""
    n, S = h["n"], h["S"]
    cand = np.ones(n, bool)
    mx = 0
    for j, col in enumerate(h["cols"]):
        sh = _shift(h, j)
        if j < S:
            mask, major = h["masks"][j], h["majors"][j]
            cout = cand & mask
            if j > 0 and not cout.any():       # the major byte is not ORed in
                cand &= ~mask
                continue
            m = int(col[cout].max()) if cout.any() else -1
            chosen = m if m > major else major
        else:
            chosen = int(col[cand].max()) if cand.any() else 0
        mx |= chosen << sh
        cand &= col == chosen
    argmax = 0
    if h["r"]:
        tc = h["tail"][cand]
        chosen = int(tc.max()) if len(tc) else 0
        mx |= chosen
        argmax = int((tc == chosen).sum())
    v = (mx + h["base"]) & ((1 << 64) - 1)
    v = v - (1 << 64) if v >= 1 << 63 else v
    return float(np.float64(np.int64(v))) / 2.0 ** h["dlen"], mx, argmax


# ---- byte-slice (buff_slice.rs:222-1970) --------------------------------------
# The stream is byte_fixed's 20-byte header followed by ceil(F/8) planes (one
# plane when F == 0) of the F-bit offset shifted left to a whole number of
# bytes, MSB plane first, every byte XOR 0x80. A filter compares the top
# min(planes, 4) bytes of that left-aligned offset with the same bytes of the
# target taken modulo 2^F (only those planes are read); the AVX2 forms cover
# the first 32*floor(n/32) rows.

def slice_planes(F):
    return max(1, -(-F // 8))


def slice_encode(x, scale):
    """Independent byte-slice encoder (uint8 stream)."""
    _, dlen = dlen_of_scale(scale)
    f = fixed(x, dlen)
    n = len(f)
    mn, mx = quirk_minmax(f)
    F = width(mn, mx)
    off = (f.astype(np.uint64) - np.uint64(mn & ((1 << 64) - 1))) if n else np.zeros(0, np.uint64)
    ns = slice_planes(F)
    pad = 8 * ns - F
    left = (off % np.uint64(1 << F) if F < 64 else off) << np.uint64(pad)
    hdr = np.zeros(20, dtype=np.uint8)
    hdr[0:8] = np.frombuffer(np.int64(mn).tobytes(), dtype=np.uint8)
    hdr[8:12] = np.frombuffer(np.uint32(n).tobytes(), dtype=np.uint8)
    hdr[12:16] = np.frombuffer(np.uint32((F - dlen) & 0xFFFFFFFF).tobytes(), dtype=np.uint8)
    hdr[16:20] = np.frombuffer(np.uint32(dlen).tobytes(), dtype=np.uint8)
    parts = [hdr]
    for k in range(ns):
        parts.append(((left >> np.uint64(8 * (ns - 1 - k))) & np.uint64(0xFF)).astype(np.uint8) ^ np.uint8(0x80))
    return np.concatenate(parts)


def slice_parse(stream):
    """Header fields and the F-bit offsets of a byte-slice stream."""
    s = np.asarray(stream, dtype=np.uint8)
    base = int(np.frombuffer(s[0:8].tobytes(), dtype=np.int64)[0])
    n = int(np.frombuffer(s[8:12].tobytes(), dtype=np.uint32)[0])
    ilen = int(np.frombuffer(s[12:16].tobytes(), dtype=np.uint32)[0])
    dlen = int(np.frombuffer(s[16:20].tobytes(), dtype=np.uint32)[0])
    F = (ilen + dlen) & 0xFFFFFFFF
    ns = slice_planes(F)
    left = np.zeros(n, dtype=np.uint64)
    for k in range(ns):
        b = s[20 + k * n: 20 + (k + 1) * n] ^ np.uint8(0x80)
        left = (left << np.uint64(8)) | b.astype(np.uint64)
    off = left >> np.uint64(8 * ns - F)
    return dict(base=base, n=n, ilen=ilen, dlen=dlen, F=F, ns=ns, off=off)


def _values(h):
    v = (np.uint64(h["base"] & ((1 << 64) - 1)) + h["off"]).view(np.int64)
    return v.astype(np.float64) / 2.0 ** h["dlen"]


def slice_decode(stream, kind="buff", rows=None):
    """buff_slice_decode; "bs" / "cond" (the four-plane readers) panic unless 1 <= F <= 32."""
    h = slice_parse(stream)
    if kind != "buff" and not 1 <= h["F"] <= 32:
        raise Panic("bit length greater than 32 is not supported yet.")
    v = _values(h)
    if kind == "cond":
        rows = np.asarray(rows, dtype=np.int64)
        if len(rows) and rows.max() >= h["n"]:
            raise Panic("chunk.get(elem).unwrap()")
        return v[rows]
    return v


def slice_max(stream, scalar=False):
    """(max value, argmax mask) of buff_slice_buffmax / buff_slice_max; the f64
    values are exact here (|base + off| < 2^53)."""
    h = slice_parse(stream)
    if scalar and not 1 <= h["F"] <= 32:
        raise Panic("bit length greater than 32 is not supported yet.")
    v = _values(h)
    if h["n"] == 0:
        return -np.finfo(np.float64).max, np.zeros(0, bool)
    return float(v.max()), v == v.max()


def slice_max_range(stream, s, e, window):
    """buff_slice_max_range: a panic unless 1 <= F <= 32 and the windows tile a
    non-empty [s, e); then every window's max and the rows holding it."""
    h = slice_parse(stream)
    if not 1 <= h["F"] <= 32:
        raise Panic("bit length greater than 32 is not supported yet.")
    L = e - s
    if window == 0 or L <= 0 or L % window:
        raise Panic("assert_eq! on the window count")
    v = _values(h)
    seg = v[s:e].reshape(-1, window)
    wm = seg.max(axis=1)
    arg = np.zeros(h["n"], bool)
    arg[s:e] = (seg == wm[:, None]).reshape(-1)
    return wm, arg


def slice_filter_bits(stream, scale, kind, pred, cond=None, semantic=False):
    """The row mask of one byte-slice filter kind (oracle.refcpu.SLICE_KINDS);
    the *_with_slice kinds take a buff_simd256 (flipped byte_fixed) stream.
    semantic: the intended comparison on every row."""
    with_slice = kind.endswith("with_slice")
    h = parse(stream, flip=True) if with_slice else slice_parse(stream)
    n, F, base = h["n"], h["F"], h["base"]
    off = h["off"]
    is_cond = kind.endswith("_cond")
    keep = np.ones(n, bool) if cond is None else np.asarray(cond, bool)
    fp = fixed_pred(pred, scale, h["dlen"])
    if fp < base:
        if semantic:
            return (np.ones(n, bool) if kind.startswith("range") else np.zeros(n, bool)) & keep
        return np.zeros(n, bool) if is_cond else np.ones(n, bool)
    t = fp - base
    if semantic:
        if F < 64 and t >= 1 << F:
            return np.zeros(n, bool) if not kind.startswith("smaller") else keep.copy()
        tt = np.uint64(t)
        res = off > tt if kind.startswith("range") else off < tt if kind.startswith("smaller") else off == tt
        return res & keep
    nosimd = kind.endswith("nosimd")
    if F == 0 or (with_slice and F < 8):
        if nosimd and n == 0:
            return np.zeros(0, bool)
        raise Panic("no slice to read")
    if not nosimd and n < 32:
        raise Panic("len - BYTE_WORD wraps")
    tm = t % (1 << F)
    if with_slice:
        P, r = F // 8, F % 8
        used = min(P, 4)
        sh = F - 8 * used
        hi_o, hi_t = off >> np.uint64(sh), np.uint64(tm >> sh)
        lo_o, lo_t = off & np.uint64((1 << r) - 1), np.uint64(tm & ((1 << r) - 1))
        if kind.startswith("range"):
            res = (hi_o > hi_t) | ((hi_o == hi_t) & (lo_o > lo_t) if r else False)
        else:
            res = (hi_o == hi_t) & (lo_o == lo_t) if r else np.zeros(n, bool)
    else:
        ns = slice_planes(F)
        used = min(ns, 4)
        sh = 8 * ns - F        # left-align, then keep the top `used` bytes
        left_o = off << np.uint64(sh)
        left_t = tm << sh
        drop = 8 * (ns - used)
        a, b = left_o >> np.uint64(drop), np.uint64(left_t >> drop)
        if kind.startswith("range"):
            res = a > b
        elif kind.startswith("smaller"):
            res = a < b
        else:
            res = a == b
    res = np.asarray(res, bool) & keep
    if not nosimd:
        res[32 * (n // 32):] = False
    return res


# --- RAPG residue format (split_double.rs:624-954, 1353-1893, 2210-2629) -------
# Written from the format, independently of oracle/refrapg.inc.

def roaring_values(b):
    """The values of a Roaring portable serialization (ascending)."""
    b = bytes(b)
    at = 0

    def u16():
        nonlocal at
        if at + 2 > len(b):
            raise ValueError("truncated")
        at += 2
        return int.from_bytes(b[at - 2: at], "little")

    def u32():
        nonlocal at
        if at + 4 > len(b):
            raise ValueError("truncated")
        at += 4
        return int.from_bytes(b[at - 4: at], "little")

    cookie = u32()
    if cookie & 0xFFFF == 12347:
        size = (cookie >> 16) + 1
        nflag = (size + 7) // 8
        flags = b[at: at + nflag]
        at += nflag
        run = [bool((flags[i // 8] >> (i % 8)) & 1) for i in range(size)]
    elif cookie == 12346:
        size = u32()
        run = [False] * size
    else:
        raise ValueError("bad cookie")
    heads = [(u16(), u16() + 1) for _ in range(size)]
    if not any(run) or size >= 4:
        at += 4 * size
    out = []
    for (key, card), r in zip(heads, run):
        hi = key << 16
        if r:
            for _ in range(u16()):
                s, l = u16(), u16()
                out.extend(hi | v for v in range(s, s + l + 1))
        elif card > 4096:
            words = np.frombuffer(b[at: at + 8192], dtype="<u8")
            at += 8192
            bits = np.unpackbits(words.view(np.uint8), bitorder="little")
            out.extend(hi | int(v) for v in np.flatnonzero(bits))
        else:
            out.extend(hi | u16() for _ in range(card))
    return out


def rapg_parse(stream):
    """Fields of a byte_residue_encode stream: header, outlier columns (rows,
    bytes, bitmap bytes), the byte after them, the dense planes and tail, and
    the offsets they add up to."""
    s = bytes(stream)
    base = int.from_bytes(s[0:8], "little", signed=True)
    n, ilen, dlen = (int.from_bytes(s[8 + 4 * i: 12 + 4 * i], "little") for i in range(3))
    F = (ilen + dlen) & 0xFFFFFFFF
    h = {"base": base, "n": n, "ilen": ilen, "dlen": dlen, "F": F, "cols": [], "S": 0}
    pos = 20
    if ilen > 15:
        S = s[20]
        h["S"] = S
        pos = 21
        for _ in range(S):
            bl = int.from_bytes(s[pos: pos + 4], "little")
            bm = s[pos + 4: pos + 4 + bl]
            pos += 4 + bl
            no = int.from_bytes(s[pos: pos + 4], "little")
            ob = np.frombuffer(s[pos + 4: pos + 4 + no], dtype=np.uint8)
            pos += 4 + no
            h["cols"].append({"bitmap": bm, "rows": np.array(roaring_values(bm), dtype=np.int64), "bytes": ob})
        if S:
            h["pad"] = s[pos]
            pos += 1
    h["planes_at"] = pos
    rem = F - 8 * h["S"]
    P, r = rem // 8, rem % 8
    h["P"], h["r"] = P, r
    h["planes"] = [np.frombuffer(s[pos + k * n: pos + (k + 1) * n], dtype=np.uint8) for k in range(P)]
    tail_at = pos + P * n
    h["tail_at"] = tail_at
    off = np.zeros(n, dtype=np.uint64)
    for j, c in enumerate(h["cols"]):
        col = np.zeros(n, dtype=np.uint64)
        col[c["rows"]] = c["bytes"]
        off |= col << np.uint64(F - 8 * (j + 1))
    for k, pl in enumerate(h["planes"]):
        off |= pl.astype(np.uint64) << np.uint64(rem - 8 * (k + 1))
    if r:
        bits = np.unpackbits(np.frombuffer(s[tail_at:], dtype=np.uint8), bitorder="little")
        idx = np.arange(n)[:, None] * r + np.arange(r)[None, :]
        off |= (bits[idx].astype(np.uint64) << np.arange(r, dtype=np.uint64)).sum(axis=1).astype(np.uint64)
    h["off"] = off
    h["stream_len"] = len(s)
    return h


def rapg_fields(stream, start, n, q):
    """n LSB-first q-bit fields of the stream from byte `start` on."""
    bits = np.unpackbits(np.frombuffer(bytes(stream)[start:], dtype=np.uint8), bitorder="little")
    if n * q > len(bits):
        raise Panic("fields run past the stream")
    idx = np.arange(n)[:, None] * q + np.arange(q)[None, :]
    return (bits[idx].astype(np.uint64) << np.arange(q, dtype=np.uint64)).sum(axis=1).astype(np.uint64)


def rapg_read(stream, precision=-1):
    """What byte_residue_decode[_with_precision] / byte_residue_sum[_with_precision]
    read: R bits (every bit for precision < 0; else the read rule), the outlier
    columns at R - 8(j+1), then whole planes, then R - 8S mod 8 = q-bit fields
    from the byte after the last plane read. Returns (h, R, read values as u64
    offsets, column u64 sums in read order, q)."""
    h = rapg_parse(stream)
    R = h["F"] if precision < 0 else precision_R(h, precision)
    S, n = h["S"], h["n"]
    rem = R - 8 * S
    k, q = rem // 8, rem % 8
    if k > h["P"]:
        raise Panic("planes run past the stream")
    v = np.zeros(n, dtype=np.uint64)
    sums = []
    for j, c in enumerate(h["cols"]):
        col = np.zeros(n, dtype=np.uint64)
        col[c["rows"]] = c["bytes"]
        v |= col << np.uint64(R - 8 * (j + 1))
        sums.append(int(c["bytes"].astype(np.uint64).sum()))
    for p in range(k):
        v |= h["planes"][p].astype(np.uint64) << np.uint64(rem - 8 * (p + 1))
        sums.append(int(h["planes"][p].astype(np.uint64).sum()))
    if q:
        f = rapg_fields(stream, h["planes_at"] + k * n, n, q)
        v |= f
        sums.append(int(f.sum()))
    return h, R, v, sums, q


def rapg_decode(stream, precision=-1):
    """byte_residue_decode[_with_precision] from the read model: (base + v) / 2^dlen
    (wrapping i64), nothing when the outlier columns take every bit read."""
    h, R, v, _, q = rapg_read(stream, precision)
    if h["S"] and R - 8 * h["S"] == 0:
        return np.zeros(0)
    f = (v + np.uint64(h["base"] & 0xFFFFFFFFFFFFFFFF)).view(np.int64)
    return f.astype(np.float64) / 2.0 ** h["dlen"]


def rapg_sum(stream, precision=-1):
    """byte_residue_sum[_with_precision] from the read model: (column sums in read
    order, exact value as a Fraction, whether the base term is added). Unlike the
    decode, the sum scales the R-bit read by 2^(ilen - R)."""
    from fractions import Fraction

    h, R, v, sums, q = rapg_read(stream, precision)
    rem = R - 8 * h["S"]
    with_base = not (rem >= 8 and rem % 8 == 0)   # the plane loop returns when the width ends on a byte
    # byte k of the read weighs 2^(ilen - bits read through it): the R-bit value v weighs 2^(ilen - R)
    sv = int(v.astype(object).sum()) if len(v) else 0
    tot = Fraction(sv) * Fraction(2) ** (h["ilen"] - R) if sv else Fraction(0)
    if with_base:
        tot += Fraction(h["n"] * h["base"], 2 ** h["dlen"])
    return sums, tot, with_base
