"""Planted mistakes in the oracle are caught by its pins.

Each case below plants one plausible slip in a copy of oracle/refcpu.cpp (a
quirk "fixed", a dropped or added term, a wrong compare, an off-by-one shift),
builds the copy into a temporary directory and runs tests/test_oracle_pins.py
against it (BUFF_ORACLE_LIB points the oracle loader at the copy). The pins
must fail on every one of them; a mutation that survives names a branch of the
oracle that nothing pins.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "refcpu.cpp")

# (name, text in refcpu.cpp, replacement, which occurrence (0-based))
MUTATIONS = [
    # Q1: the `else if` running max of byte_fixed_encode (split_double.rs:321-330) made an `if`
    ("encode_if_for_else_if", "else if (f > mx) mx = f;", "if (f > mx) mx = f;", 0),
    # Q6: byte_fixed_sum's early return (split_double.rs:2171-2173) given the base term it drops
    ("sum_first_return_adds_base", "if (remain == 0) return sum;   // returns without the base term",
     "if (remain == 0) return sum + sum_base;", 0),
    # the same, in the loop over the later sub-columns (split_double.rs:2188-2190)
    ("sum_loop_return_adds_base", "    if (remain == 0) return sum;\n  }\n  if (remain > 0) {\n    if (processed >= 8) bp.finish_read_byte();\n    processed += (int32_t)remain;\n    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));\n    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);\n    if (plane_sums) plane_sums[k] = sum_fixed;",
     "    if (remain == 0) return sum + sum_base;\n  }\n  if (remain > 0) {\n    if (processed >= 8) bp.finish_read_byte();\n    processed += (int32_t)remain;\n    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));\n    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);\n    if (plane_sums) plane_sums[k] = sum_fixed;", 0),
    # byte_fixed_sum's tail term weighed as a whole byte
    ("sum_tail_weight", "    processed += (int32_t)remain;\n    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));\n    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);\n    if (plane_sums) plane_sums[k] = sum_fixed;",
     "    processed += 8;\n    dec_scl = powi(2.0, (int32_t)((uint32_t)processed - h.ilen));\n    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);\n    if (plane_sums) plane_sums[k] = sum_fixed;", 0),
    # byte_residue_sum_with_precision's early return (split_double.rs:2590-2592, 2607-2609) given the base term
    ("psum_return_adds_base", "    if (remain == 0) return sum;\n  }\n  if (remain > 0) {   // 2611-2622",
     "    if (remain == 0) return sum + sum_base;\n  }\n  if (remain > 0) {   // 2611-2622", 0),
    # the precision read rule's round-up to a whole byte (split_double.rs:2455-2460) dropped in the sum
    ("psum_no_roundup", "if (byte_needed < fixed_byte && extra_bits > 0) { byte_needed += 1; remain = byte_needed * 8; }",
     "if (false) { byte_needed += 1; remain = byte_needed * 8; }", 1),
    # the q-bit fields of the precision sum weighed as a whole byte (split_double.rs:2615-2616)
    ("psum_fields_weight", "  if (remain > 0) {   // 2611-2622\n    if (processed >= 8) bp.finish_read_byte();\n    processed += (int32_t)remain;",
     "  if (remain > 0) {   // 2611-2622\n    if (processed >= 8) bp.finish_read_byte();\n    processed += 8;", 0),
    # the precision sum's fields read at the stored tail width instead of q bits
    ("psum_fields_width", "    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits(remain);\n    if (plane_sums && k < 9) plane_sums[k] = sum_fixed;",
     "    for (uint32_t i = 0; i < h.len; ++i) sum_fixed += bp.read_bits((h.ilen + h.dlen) % 8 ? (h.ilen + h.dlen) % 8 : remain);\n    if (plane_sums && k < 9) plane_sums[k] = sum_fixed;", 0),
    # range filter: `>` on the top sub-column made `>=` (buff_simd.rs:4546-4561)
    ("range_ge_top_plane", "if (c > cur_tar) res.set(i);", "if (c >= cur_tar) res.set(i);", 0),
    # range filter: the tail target not masked to the low r bits (buff_simd.rs:4632)
    ("range_tail_target_unmasked", "cur_tar = (uint8_t)((uint8_t)((uint8_t)fixed_target << sh) >> sh);",
     "cur_tar = (uint8_t)fixed_target;", 0),
    # equality filter: `==` on the top sub-column made `>=` (buff_simd.rs:4966-4970)
    ("equal_ge_top_plane", "if (chunk[i] == dec_byte) rb1.set(i);", "if (chunk[i] >= dec_byte) rb1.set(i);", 0),
    # decode: the top sub-column shifted one bit too far (split_double.rs:1300-1310)
    ("decode_plane_shift", "for (uint32_t i = 0; i < h.len; ++i) fixed_vec.push_back((uint64_t)F(chunk[i]) << (remain & 63));",
     "for (uint32_t i = 0; i < h.len; ++i) fixed_vec.push_back((uint64_t)F(chunk[i]) << ((remain + 1) & 63));", 0),
    # max: the base left out of the printed value (split_double.rs:4102)
    ("max_without_base", "return (double)(int64_t)(mx + (uint64_t)h.base) / powi(2.0, (int32_t)h.dlen);",
     "return (double)(int64_t)(mx) / powi(2.0, (int32_t)h.dlen);", 0),
]


# mistakes planted in the RAPG restatement (oracle/refrapg.inc), pinned by tests/test_rapg_oracle.py
RAPG_MUTATIONS = [
    # CRoaring's run_optimize keeps an array whose run form is the same size (2 + 4 runs == 2 card + 2)
    ("roaring_run_ties_become_runs", "if ((c.type == 0 && as_run < 2 * card + 2)",
     "if ((c.type == 0 && as_run <= 2 * card + 2)", 0),
    # the offset table is also written with runs once there are four or more containers
    ("roaring_no_offsets_with_runs", "if (!hasrun || size >= kNoOffsetThreshold) {\n    for (auto& c : cs) {",
     "if (!hasrun) {\n    for (auto& c : cs) {", 0),
    # byte_residue_encode's min / max is `if` / `if` (split_double.rs:642-651), not byte_fixed's `else if`
    ("residue_encode_else_if", "    if (f > mx) mx = f;\n    fixed_vec.push_back(f);",
     "    else if (f > mx) mx = f;\n    fixed_vec.push_back(f);", 0),
    # the sum's early return on a byte boundary (split_double.rs:2271-2273) given the base term
    ("residue_sum_return_adds_base", "if (returned) return s;", "if (false) return s;", 0),
    # an outlier column's byte ORed one byte too high (split_double.rs:1411)
    ("residue_decode_outlier_shift", "fixed_vec[i] |= (uint64_t)outlier_value << (remain & 63);",
     "fixed_vec[i] |= (uint64_t)outlier_value << ((remain + 8) & 63);", 0),
]


def _mutate(src, old, new, occ):
    pos = -1
    for _ in range(occ + 1):
        pos = src.find(old, pos + 1)
        if pos < 0:
            raise AssertionError("mutation anchor not found: %r" % old[:60])
    return src[:pos] + new + src[pos + len(old):]


def _build_and_pin(tmp, name, old, new, occ, target=SRC, pins="test_oracle_pins.py"):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    cpp = os.path.join(d, "refcpu.cpp")
    with open(SRC) as f:
        src = f.read()
    if target == SRC:
        src = _mutate(src, old, new, occ)
    else:   # an included file: the mutated copy sits beside the copy that includes it
        with open(target) as f:
            inc = f.read()
        with open(os.path.join(d, os.path.basename(target)), "w") as f:
            f.write(_mutate(inc, old, new, occ))
    with open(cpp, "w") as f:
        f.write(src)
    # the mutated copy includes the oracle's other sources from their own directory
    so = os.path.join(d, "librefcpu.so")
    b = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fno-fast-math", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "oracle"), "-o", so, cpp], capture_output=True, text=True)
    if b.returncode != 0:
        return name, "build failed: " + b.stderr[-2000:]
    env = dict(os.environ, BUFF_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", pins), "-x",
                        "-q", "-p", "no:cacheprovider", "-m", "not gpu"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=600)
    return name, (r.returncode, r.stdout[-3000:])


RAPG_SRC = os.path.join(ROOT, "oracle", "refrapg.inc")


def test_mutations_are_caught():
    jobs = [m + (SRC, "test_oracle_pins.py") for m in MUTATIONS] + \
           [m + (RAPG_SRC, "test_rapg_oracle.py") for m in RAPG_MUTATIONS]
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(4) as ex:
            results = list(ex.map(lambda m: _build_and_pin(tmp, *m), jobs))
    survivors = []
    for name, res in results:
        assert not isinstance(res, str), (name, res)
        rc, out = res
        # a pin failed (pytest exit 1), not a collection or usage error
        if rc != 1 or " failed" not in out:
            survivors.append((name, rc, out[-600:]))
    assert not survivors, survivors


@pytest.mark.parametrize("name,old,new,occ,src", [m + (SRC,) for m in MUTATIONS] +
                         [m + (RAPG_SRC,) for m in RAPG_MUTATIONS])
def test_mutation_anchor_present(name, old, new, occ, src):
    with open(src) as f:
        _mutate(f.read(), old, new, occ)
