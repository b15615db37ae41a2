"""The reference README quick-start: its input file, scale 10000, pred 1.1509.

`cargo run --bin comp_profiler ./data/randomwalkdatasample1k-1k buff-simd 10000 1.1509`
(README.md:12). The reference stores no outputs for it, so the CPU tests pin
the oracle's results with the closed forms, and the GPU tests run the same
input through the C ABI and through the comp_profiler CLI.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests import closed_form as cf
from tests import refdata

SCALE, PRED = 10000, 1.1509


@pytest.fixture(scope="module")
def data():
    x = refdata.quickstart_values()
    assert len(x) == 500_000
    return x


def test_ingest_matches_fixture_shape(data):
    assert data.min() == -119.11 and data.max() == 106.9


def test_oracle_on_quickstart(oracle, data):
    s = oracle.encode(data, SCALE)
    assert np.array_equal(s, cf.encode(data, SCALE))
    for op in ("range", "equal"):
        w, c, _ = oracle.buff_filter(s, SCALE, op, PRED)
        m = cf.filter_bits(s, SCALE, op, PRED)
        assert np.array_equal(w, cf.to_words(m)) and c == int(m.sum())
    h = cf.parse(s)
    assert oracle.byte_fixed_max(s, SCALE)[1] == int(h["off"].max())


@pytest.mark.gpu
def test_gpu_quickstart(gpu_ctx, oracle, data):
    s = oracle.encode(data, SCALE)
    ch = gpu_ctx.encode(data, SCALE)
    assert np.array_equal(ch.serialize(), s)
    for op in ("range", "equal"):
        w_ref, c_ref, _ = oracle.buff_filter(s, SCALE, op, PRED)
        w, c = ch.filter(op, PRED)
        assert c == c_ref and np.array_equal(w, w_ref)
    assert ch.sum()[0] == oracle.byte_fixed_sum(s, SCALE)[0]
    assert ch.max()[0] == oracle.byte_fixed_max(s, SCALE)[0]
    ch.free()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["buff", "buff-simd", "buff-slice", "buff-slice-scalar", "buff-major"])
def test_comp_profiler_cli(tmp_path, oracle, data, method):
    """The C++ host mirror + CLI prints what the reference driver prints."""
    from tests.conftest import ROOT

    exe = os.path.join(ROOT, "buff_amd", "comp_profiler")
    assert os.path.exists(exe), "comp_profiler is built by buff_amd.build"
    csv = tmp_path / "randomwalkdatasample1k-1k"
    csv.write_text(refdata.quickstart_csv_text())
    out = subprocess.run([exe, str(csv), method, str(SCALE), str(PRED)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    s = oracle.encode(data, SCALE)
    perf = [ln for ln in out.stdout.splitlines() if ln.startswith("Performance:")]
    assert len(perf) == 1 and len(perf[0].split(",")) == 10
    if method.startswith("buff-slice"):
        sl = oracle.slice_encode(data, SCALE)
        kinds = ("range_nosimd", "equal_nosimd") if method.endswith("scalar") else ("range", "equal")
        for kind in kinds:
            assert "Number of qualified items:%d" % oracle.slice_filter(sl, SCALE, kind, PRED)[1] in out.stdout
        mx, cnt, _ = oracle.slice_max(sl, method.endswith("scalar"))
        assert "Number of qualified max items:%d" % cnt in out.stdout
        m = re.search(r"Max value:(\S+)", out.stdout)
        assert m and float(m.group(1)) == mx
        m = re.search(r"sum is:(\S+)", out.stdout)
        vals = oracle.slice_decode(sl)
        assert m and abs(float(m.group(1)) - oracle.slice_sum(sl)) <= 1e-12 * np.abs(vals).sum()
        assert abs(float(perf[0].split(",")[3]) - len(sl) / (31 + 8 * len(data))) < 1e-12
        return
    if method == "buff-major":
        mj = oracle.major_encode(data, SCALE)
        assert "Number of qualified items:%d" % oracle.major_filter(mj, SCALE, "range", PRED)[1] in out.stdout
        assert abs(float(perf[0].split(",")[3]) - len(mj) / (31 + 8 * len(data))) < 1e-12
        return
    c_range = oracle.buff_filter(s, SCALE, "range", PRED)[1]
    c_eq = oracle.buff_filter(s, SCALE, "equal", PRED)[1]
    assert "Number of qualified items:%d" % c_range in out.stdout
    assert "Number of qualified int items:%d" % c_eq in out.stdout
    if method == "buff-simd":
        wsl = oracle.slice_filter(oracle.encode(data, SCALE, flip=True), SCALE, "range_with_slice", PRED)[1]
        assert "Number of qualified int items:%d" % wsl in out.stdout
    if method == "buff":
        m = re.search(r"sum is: (\S+)", out.stdout)
        assert m and float(m.group(1)) == oracle.byte_fixed_sum(s, SCALE)[0]
        m = re.search(r"Max value:(\S+)", out.stdout)
        assert m and float(m.group(1)) == oracle.byte_fixed_max(s, SCALE)[0]
    perf = [ln for ln in out.stdout.splitlines() if ln.startswith("Performance:")]
    assert len(perf) == 1
    fields = perf[0].split(",")
    assert len(fields) == 10
    ratio = float(fields[3])
    assert abs(ratio - len(s) / (31 + 8 * len(data))) < 1e-12
