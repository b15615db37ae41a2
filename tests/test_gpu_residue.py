"""The RAPG residue format on the GPU (BUFF_RESIDUE) against the oracle.

Every column of the oracle's residue cases is encoded on the GPU: the stream
must equal byte_residue_encode's (split_double.rs:624-954) byte for byte,
CRoaring bitmaps included; the oracle's stream loaded back must serialize to
itself. byte_residue_decode and byte_residue_decode_with_precision
(1353-1893) must return the oracle's values bit for bit, and byte_residue_sum
and byte_residue_sum_with_precision (2210-2629) its f64 sum bit for bit with
its u64 column sums, for p = 0..6 and the full read; inputs on which the
reference panics must report BUFF_E_REF_PANIC. The one read the GPU does not
reproduce — a reduced-precision decode of a stream with outlier columns whose
read ends in fields of the packed tail — must be exactly where the read model
puts it, and report BUFF_E_UNSUPPORTED. Filters and maxima answer only with
BUFF_SEMANTIC (the true answer on the offsets).
"""
import numpy as np
import pytest

from buff_amd import _native as N
from buff_amd import datagen as dg
from buff_amd.split_double import SplitBDDoubleCompress
from tests import closed_form as cf
from tests.test_rapg_oracle import CASES

pytestmark = pytest.mark.gpu

NAMES = sorted(CASES)


@pytest.fixture(scope="module")
def streams(oracle):
    return {name: oracle.residue_encode(*CASES[name])[0] for name in NAMES}


@pytest.mark.parametrize("name", NAMES)
def test_encode_and_load_bit_exact(gpu_ctx, streams, name):
    x, scale = CASES[name]
    ref = streams[name]
    print("case", name, len(x), flush=True)
    enc = gpu_ctx.encode(x, scale, flags=N.BUFF_RESIDUE)
    loaded = gpu_ctx.load(ref, scale, flags=N.BUFF_RESIDUE)
    try:
        for ch in (enc, loaded):
            got = ch.serialize()
            assert len(got) == len(ref), name
            assert np.array_equal(got, ref), (name, int(np.argmax(got != ref)))
    finally:
        enc.free()
        loaded.free()


def _expect_unsupported(stream, p):
    """The reads the GPU leaves out: outlier columns with a read ending in fields of the packed tail."""
    h, R, _, _, q = cf.rapg_read(stream, p)
    k = (R - 8 * h["S"]) // 8
    return h["S"] > 0 and 0 < q != h["r"] and k == h["P"]


@pytest.mark.parametrize("name", NAMES)
def test_decode_and_sum_match_oracle(oracle, streams, name):
    x, scale = CASES[name]
    s = streams[name]
    sd = SplitBDDoubleCompress(1, 1, scale)
    print("case", name, len(x), flush=True)
    for p in [-1, 0, 1, 2, 3, 4, 5, 6]:
        try:
            want = oracle.residue_decode(s, p)
        except oracle.RefPanic:
            with pytest.raises(N.RefPanic):
                sd.byte_residue_decode(s) if p < 0 else sd.byte_residue_decode_with_precision(s, p)
            want = None
        if want is not None:
            if p >= 0 and _expect_unsupported(s, p):
                with pytest.raises(N.Unsupported):
                    sd.byte_residue_decode_with_precision(s, p)
            else:
                got = sd.byte_residue_decode(s) if p < 0 else sd.byte_residue_decode_with_precision(s, p)
                assert len(got) == len(want), (name, p)
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, p)
        try:
            s_ref, ps_ref, _ = oracle.residue_sum(s, p)
        except oracle.RefPanic:
            with pytest.raises(N.RefPanic):
                sd.byte_residue_sum(s) if p < 0 else sd.byte_residue_sum_with_precision(s, p)
            continue
        got = sd.byte_residue_sum(s) if p < 0 else sd.byte_residue_sum_with_precision(s, p)
        assert got == s_ref or (np.isnan(got) and np.isnan(s_ref)), (name, p, got, s_ref)


@pytest.mark.parametrize("name", ["spikes_5pct", "spikes_p5", "walk", "outlier_heavy"])
def test_chunk_sums_and_semantic_answers(gpu_ctx, oracle, streams, name):
    x, scale = CASES[name]
    s = streams[name]
    dlen = cf.dlen_of_scale(scale)[1]
    f = cf.fixed(x, dlen)
    ch = gpu_ctx.load(s, scale, flags=N.BUFF_RESIDUE)
    try:
        val, exact, sums = ch.sum()
        s_ref, ps_ref, _ = oracle.residue_sum(s)
        assert val == s_ref and np.array_equal(sums[: 9], ps_ref[: 9]), name
        tot = int(f.astype(object).sum())
        assert abs(exact - tot / 2.0 ** dlen) <= 1e-12 * max(1.0, abs(tot / 2.0 ** dlen))
        # no reference filter / max of a residue stream: the true answers with BUFF_SEMANTIC
        with pytest.raises(N.Unsupported):
            ch.filter("range", float(np.median(x)))
        pred = float(np.quantile(x, 0.9, method="lower"))
        w, c = ch.filter("range", pred, flags=N.BUFF_SEMANTIC)
        m = f > cf.fixed(np.array([pred]), dlen)[0]
        assert c == int(m.sum()) and np.array_equal(w, cf.to_words(m)), name
        v, mo, _ = ch.max_ex(flags=N.BUFF_SEMANTIC)
        assert mo == int(f.max() - f.min()), name
        d = ch.materialize(-1, N.BUFF_SEMANTIC)
        assert np.array_equal(d, f / 2.0 ** dlen)
    finally:
        ch.free()


def test_residue_4m_rows(gpu_ctx, oracle):
    """4,194,304 rows with 1% spikes: the stream, the full decode against the
    fixed grid, the sums and the reduced-precision reads against the oracle."""
    x = dg.positive_spikes(1 << 22, 3, 31, 0.01)
    s, cols = oracle.residue_encode(x, 1000)
    assert cols >= 1
    ch = gpu_ctx.encode(x, 1000, flags=N.BUFF_RESIDUE)
    try:
        assert np.array_equal(ch.serialize(), s)
        d = ch.materialize()
        assert np.array_equal(d, cf.fixed(x, 11) / 2.0 ** 11)
        val, _, _ = ch.sum()
        assert val == oracle.residue_sum(s)[0]
        for p in (0, 1, 2, 3):
            r, _, _ = ch.sum_precision(p)
            assert r == oracle.residue_sum(s, p)[0], p
    finally:
        ch.free()


def test_encode_panics_and_limits(gpu_ctx):
    with pytest.raises(N.RefPanic):   # ilen > 15 samples the first 2000 offsets
        gpu_ctx.encode(dg.positive_spikes(1999, 3, 9), 1000, flags=N.BUFF_RESIDUE)
    x = np.zeros(5000)
    x[4000] = 2.0 ** 40   # the leading-zero histogram read past its 32 slots
    with pytest.raises(N.RefPanic):
        gpu_ctx.encode(x, 1000, flags=N.BUFF_RESIDUE)
    with pytest.raises(N.BuffError):
        gpu_ctx.encode(x, 1000, flags=N.BUFF_RESIDUE | N.BUFF_MAJORITY)


def test_load_rejects_a_run_past_its_container(gpu_ctx, oracle):
    """A CRoaring run container whose run leaves its 65536-row container (the
    encoder never writes one) is a malformed stream on load, not rows of the
    next container."""
    x = np.full(70000, 100.0)
    x[1000:1100] = 1e6 + np.arange(100)
    s, cols = oracle.residue_encode(x, 1000)
    h = cf.rapg_parse(s)
    assert cols == 2 and h["S"] == 2 and list(h["cols"][0]["rows"]) == list(range(1000, 1100))
    b = bytearray(s)
    # column 0's bitmap at 25: cookie 12347 (one container), run flags, key 0 /
    # card - 1, then the run container: count 1, start 1000, length - 1 = 99
    assert bytes(b[25:40]).hex() == "3b30000001000063000100e8036300"
    ch = gpu_ctx.load(np.frombuffer(bytes(b), dtype=np.uint8), 1000, flags=N.BUFF_RESIDUE)
    ch.free()
    b[38:40] = (0xFFFF).to_bytes(2, "little")   # the run now reaches row 1000 + 65535
    with pytest.raises(N.BuffError) as e:
        gpu_ctx.load(np.frombuffer(bytes(b), dtype=np.uint8), 1000, flags=N.BUFF_RESIDUE)
    assert e.value.code == N.BUFF_E_INVALID, str(e.value)
