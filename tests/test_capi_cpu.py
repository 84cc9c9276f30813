"""The C ABI library loads and exports every entry point include/wos_hip.h
declares (no compute: there is no GPU here); the ctypes mirror matches the
header's struct layouts."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from wos_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wos_hip.h")


RUN_HEADER = os.path.join(ROOT, "include", "wos_run.h")
RUN_LIB = os.path.join(ROOT, "walk-on-spheres_amd", "csrc", "libwos_run.so")


def declared_functions(header=HEADER):
    text = open(header).read()
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


def test_run_library_exports_the_registry_entry_point():
    """libwos_run.so exports what include/wos_run.h declares, and its
    registry lists the reference's equations in the reference's order."""
    if not os.path.exists(RUN_LIB):
        pytest.fail("libwos_run.so is not built; run __graft_entry__.build()")
    lib = C.CDLL(RUN_LIB)
    names = declared_functions(RUN_HEADER)
    assert "wos_hip_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.wos_hip_equation_name.restype = C.c_char_p
    n = lib.wos_hip_equation_count()
    assert [lib.wos_hip_equation_name(i) for i in range(n)] == [b"laplace", b"poisson",
                                                                  b"screened_poisson",
                                                                  b"helmholtz"]
    assert lib.wos_hip_equation_name(n) is None
    # argument errors return 1 before any device use
    lib.wos_hip_run.restype = C.c_int
    assert lib.wos_hip_run(4, b"m.obj", b"o.h5", 2, 2, 1, 1, C.c_double(1e-3), 10, 12,
                           C.c_uint64(1), C.c_double(5.0), 0, 0, None, None) == 1
    assert lib.wos_hip_run(2, b"m.obj", b"o.h5", 2, 2, 1, 1, C.c_double(1e-3), 10, 12,
                           C.c_uint64(1), C.c_double(5.0), 0, 7, None, None) == 1


def test_version_and_error_string_without_gpu():
    lib = capi.load()
    assert b"gfx950" in lib.wos_version()
    assert isinstance(lib.wos_last_error(), bytes)


STRUCTS = [("wos_eq_desc", capi.EqDesc), ("wos_run_params", capi.RunParams),
           ("wos_point_stats", capi.PointStats), ("wos_run_summary", capi.RunSummary),
           ("wos_tuning", capi.Tuning), ("wos_field_source", capi.FieldSource),
           ("wos_heat_term", capi.HeatTerm)]


def test_struct_layouts(tmp_path):
    """The ctypes mirror has the C compiler's size and field offsets for
    every structure of the header."""
    import subprocess
    src = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS:
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    src += ["return 0;", "}"]
    c_file = tmp_path / "layout.c"
    c_file.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(c_file)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in STRUCTS:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_null_context_is_an_input_error():
    lib = capi.load()
    rc = lib.wos_ctx_set_mesh(None, 2, None, 0, None, 0, None)
    assert rc == capi.WOS_ERR_INPUT
    assert b"NULL" in lib.wos_last_error()


def _host_bvh(lib, V, F):
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    dim = F.shape[1]
    n = lib.wos_bvh_build_host(dim, capi.ptr(V, C.c_double), len(V), capi.ptr(F, C.c_int32),
                               len(F), None, None, None, None, None)
    assert n > 0
    cnt = np.zeros(n, np.int32)
    idx = np.zeros(n, np.int32)
    lo = np.zeros((n, 3))
    hi = np.zeros((n, 3))
    slots = np.zeros(len(F), np.int32)
    lib.wos_bvh_build_host(dim, capi.ptr(V, C.c_double), len(V), capi.ptr(F, C.c_int32), len(F),
                           capi.ptr(cnt, C.c_int32), capi.ptr(idx, C.c_int32),
                           capi.ptr(lo, C.c_double), capi.ptr(hi, C.c_double),
                           capi.ptr(slots, C.c_int32))
    return cnt, idx, lo, hi, slots


def test_host_bvh_matches_reference_golden():
    """The product's BVH builder reproduces the reference node array and its
    leaf-slot order (same split rule, std::nth_element with the same order)."""
    lib = capi.load()
    g = np.load(os.path.join(ROOT, "tests", "golden", "bvh.npz"), allow_pickle=False)
    names = sorted({k.split("__")[0] for k in g.files})
    for name in names:
        cnt, idx, lo, hi, slots = _host_bvh(lib, g[f"{name}__mesh_verts"], g[f"{name}__mesh_prims"])
        assert np.array_equal(cnt, g[f"{name}__prim_count"]), name
        assert np.array_equal(idx, g[f"{name}__idx"]), name
        assert np.array_equal(lo, g[f"{name}__bmin"]), name
        assert np.array_equal(hi, g[f"{name}__bmax"]), name
        assert np.array_equal(slots, g[f"{name}__slots"]), name


def test_host_bvh_rejects_bad_mesh():
    lib = capi.load()
    V = np.zeros((3, 2))
    F = np.array([[0, 7]], np.int32)
    assert lib.wos_bvh_build_host(2, capi.ptr(V, C.c_double), 3, capi.ptr(F, C.c_int32), 1,
                                  None, None, None, None, None) == -1
    assert b"out of range" in lib.wos_last_error()
