"""The Rust binding crate (buff_amd/rust) against include/buffgpu.h, on the CPU.

No Rust toolchain is installed here, so the crate is checked as text: every
entry point the header declares is in the crate's extern "C" block with the
same parameters in the same order and the Rust type each C type maps to; every
#define has a constant of the same value; the #[repr(C)] structs list the
header's fields in the header's order; and every function the safe wrapper
calls is declared.
"""
import os
import re

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "buffgpu.h")
FFI = os.path.join(ROOT, "buff_amd", "rust", "src", "ffi.rs")
LIB = os.path.join(ROOT, "buff_amd", "rust", "src", "lib.rs")

# C parameter type -> the Rust FFI type the crate must use
C_TO_RUST = {
    "buff_ctx*": "*mut BuffCtx", "buff_ctx**": "*mut *mut BuffCtx",
    "const buff_chunk*": "*const BuffChunk", "buff_chunk*": "*mut BuffChunk", "buff_chunk**": "*mut *mut BuffChunk",
    "buff_chunk* const*": "*const *mut BuffChunk",
    "buff_batch*": "*mut BuffBatch", "const buff_batch*": "*const BuffBatch", "buff_batch**": "*mut *mut BuffBatch",
    "buff_server*": "*mut BuffServer", "buff_server**": "*mut *mut BuffServer",
    "const buff_server_cfg*": "*const BuffServerCfg",
    "const buff_query*": "*const BuffQuery", "buff_query_result*": "*mut BuffQueryResult",
    "buff_chunk_info*": "*mut BuffChunkInfo", "buff_tune_choice*": "*mut BuffTuneChoice",
    "const buff_split_double*": "*const BuffSplitDouble", "const buff_slice_compress*": "*const BuffSplitDouble",
    "const buff_segment*": "*const BuffSegment", "buff_segment*": "*mut BuffSegment",
    "const double*": "*const f64", "double*": "*mut f64", "double": "f64",
    "const uint8_t*": "*const u8", "uint8_t*": "*mut u8", "uint8_t": "u8",
    "const uint16_t*": "*const u16", "const uint32_t*": "*const u32", "uint32_t*": "*mut u32", "uint32_t": "u32",
    "const uint64_t*": "*const u64", "uint64_t*": "*mut u64", "uint64_t": "u64",
    "const int64_t*": "*const i64", "int64_t*": "*mut i64", "int64_t": "i64",
    "int32_t": "i32", "int32_t*": "*mut i32", "int": "c_int", "int*": "*mut c_int",
    "void*": "*mut c_void", "const void*": "*const c_void", "void* const*": "*const *mut c_void",
    "const void* const*": "*const *const c_void", "void**": "*mut *mut c_void",
    "const uint64_t**": "*mut *const u64", "const double**": "*mut *const f64",
    "const char*": "*const c_char", "float*": "*mut f32",
}


def _strip(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _header_protos():
    text = _strip(open(HEADER).read())
    out = {}
    for ret, name, params in re.findall(r"^\s*(int|const char\*)\s+(buff_\w+)\s*\(([^;]*?)\)\s*;", text,
                                        flags=re.M | re.S):
        types = []
        for p in params.replace("\n", " ").split(","):
            p = p.strip()
            if not p or p == "void":
                continue
            ty = re.match(r"^(.*?)(\w+)$", p).group(1).strip()
            ty = re.sub(r"\s*\*", "*", ty).replace("*const", "* const")
            types.append(re.sub(r"\s+", " ", ty))
        out[name] = ("c_int" if ret == "int" else "*const c_char", types)
    return out


def _rust_decls():
    text = open(FFI).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for name, params, ret in re.findall(r"pub fn (buff_\w+)\((.*?)\)\s*->\s*([^;]+);", block, flags=re.S):
        types = [p.split(":", 1)[1].strip() for p in params.split(",") if p.strip()]
        out[name] = (ret.strip(), types)
    return out


def test_every_entry_point_is_declared_with_its_types():
    header, rust = _header_protos(), _rust_decls()
    assert len(header) >= 90
    assert set(header) == set(rust), (set(header) ^ set(rust))
    for name, (ret, ctypes_) in header.items():
        rret, rtypes = rust[name]
        assert rret == ret, name
        assert len(rtypes) == len(ctypes_), name
        for i, (c, r) in enumerate(zip(ctypes_, rtypes)):
            assert c in C_TO_RUST, (name, i, c)
            assert C_TO_RUST[c] == r, (name, i, c, r)


def test_constants_match_the_header():
    text = open(HEADER).read()
    defines = dict(re.findall(r"^#define\s+(BUFF_\w+)\s+(0x[0-9a-fA-F]+|\d+)u?\b", text, flags=re.M))
    rust = dict(re.findall(r"pub const (BUFF_\w+): \w+ = (0x[0-9a-fA-F]+|\d+);", open(FFI).read()))
    assert defines and set(defines) == set(rust), set(defines) ^ set(rust)
    for k, v in defines.items():
        assert int(v, 0) == int(rust[k], 0), k


def _header_struct_fields(tag):
    text = _strip(open(HEADER).read())
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + tag + r"\s*;", text)
    assert m, tag
    return re.findall(r"(\w+)\s*;", m.group(1))


def _rust_struct_fields(name):
    text = open(FFI).read()
    m = re.search(r"pub struct " + name + r" \{(.*?)\n\}", text, flags=re.S)
    assert m, name
    return re.findall(r"pub (\w+):", m.group(1))


def test_structs_follow_the_header_layout():
    for c, r in (("buff_chunk_info", "BuffChunkInfo"), ("buff_tune_choice", "BuffTuneChoice"),
                 ("buff_query", "BuffQuery"), ("buff_query_result", "BuffQueryResult"),
                 ("buff_server_cfg", "BuffServerCfg"), ("buff_split_double", "BuffSplitDouble"),
                 ("buff_segment", "BuffSegment")):
        assert _header_struct_fields(c) == _rust_struct_fields(r), c
    assert "#[repr(C)]" in open(FFI).read()


def test_wrapper_calls_declared_functions():
    used = set(re.findall(r"ffi::(buff_\w+)", open(LIB).read()))
    declared = set(_rust_decls())
    assert used and used <= declared, used - declared
    # the reference-shaped methods the drivers call are all wrapped
    for m in ("byte_fixed_encode", "byte_fixed_decode", "buff_range_filter", "buff_equal_filter",
              "byte_fixed_sum", "byte_fixed_max", "byte_residue_decode_with_precision",
              "byte_residue_sum_with_precision", "buff_encode_majority", "buff_slice_encode"):
        assert re.search(r"pub fn " + m + r"\(", open(LIB).read()), m
