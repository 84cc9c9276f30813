"""The C ABI library loads and exports every entry point include/wos_hip.h
declares (no compute: there is no GPU here); the ctypes mirror matches the
header's struct layouts."""
import ctypes as C
import os
import re

import pytest

from wos_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wos_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(wos_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_lists_entry_points():
    names = declared_functions()
    for n in ("wos_ctx_create", "wos_ctx_destroy", "wos_ctx_set_mesh", "wos_solve_grid",
              "wos_solve_points", "wos_last_error", "wos_version"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libwos_hip.so is not built; run __graft_entry__.build()")
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    lib = capi.load()
    assert b"gfx950" in lib.wos_version()
    assert isinstance(lib.wos_last_error(), bytes)


def test_struct_layouts():
    # sizes follow the C declaration order with natural alignment
    assert C.sizeof(capi.EqDesc) == 4 * 4 + 8 * 6 + 8 * capi.WOS_MAX_BND_VALUES + 8 + 8 + 8
    assert C.sizeof(capi.RunParams) == 4 * 4 + 8 + 8 + 8 + 4 + 4
    assert C.sizeof(capi.PointStats) == 8 * 8
    assert capi.EqDesc.alpha.offset == 16
    assert capi.RunParams.epsilon.offset == 16


def test_null_context_is_an_input_error():
    lib = capi.load()
    rc = lib.wos_ctx_set_mesh(None, 2, None, 0, None, 0, None)
    assert rc == capi.WOS_ERR_INPUT
    assert b"NULL" in lib.wos_last_error()
