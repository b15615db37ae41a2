"""The C ABI library loads on any host and exports every declared symbol.

No compute call is made: these tests run without a GPU.
"""
import os
import re

from tests.conftest import ROOT


def _declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int|const char\*)\s+(buff_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 30
    for must in ("buff_encode_device", "buff_filter_device", "buff_sum", "buff_max", "buff_materialize",
                 "buff_chunk_serialize", "buff_chunk_load", "buff_sd_byte_fixed_encode",
                 "buff_sd_buff_range_filter", "buff_sd_byte_fixed_sum"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import buff_amd

    L = buff_amd.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from buff_amd import _native

    declared = set(_declared())
    bound = set(_native.SIGNATURES) | {"buff_last_error"}
    assert declared <= bound, declared - bound


def test_binding_constants_match_the_header():
    # every flag, op and status code the header defines has the same value in
    # the Python binding (a mismatch would pass the wrong bits through the ABI)
    from buff_amd import _native as N

    text = open(os.path.join(ROOT, "include", "buffgpu.h")).read()
    defines = dict(re.findall(r"^#define\s+(BUFF_\w+)\s+(0x[0-9a-fA-F]+|\d+)u?\b", text, flags=re.M))
    assert {"BUFF_SLICE", "BUFF_SCALAR", "BUFF_WITH_SLICE", "BUFF_OP_LT", "BUFF_E_REF_PANIC"} <= set(defines)
    for name, value in defines.items():
        if hasattr(N, name):
            assert getattr(N, name) == int(value, 0), name
    for name in ("BUFF_FLIP", "BUFF_SEMANTIC", "BUFF_MAJORITY", "BUFF_SLICE", "BUFF_SCALAR", "BUFF_WITH_SLICE",
                 "BUFF_OP_GT", "BUFF_OP_EQ", "BUFF_OP_LT"):
        assert hasattr(N, name), name
