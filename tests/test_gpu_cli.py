"""The `wos` tool end to end on the GPU: progress/summary lines, the HDF5
layout of include/wos/hdf5_io.hpp:147-322 (read back with h5dump), and the
stored fields equal to the library's own solve with the same seed."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WOS = os.path.join(ROOT, "walk-on-spheres_amd", "bin", "wos")
H5DUMP = "/opt/conda/bin/h5dump"


def h5_array(path, name, dtype, tmp):
    out = os.path.join(tmp, name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True,
                   capture_output=True, timeout=60)
    return np.fromfile(out, dtype=dtype)


@pytest.mark.parametrize("dim", [2, 3])
def test_cli_writes_reference_layout(tmp_path, dim):
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    if dim == 2:
        V, F, B = M.annulus(64)
        args = ["laplace", "12", "10"]
        grid = (12, 10, 1)
        eq = Pr.laplace_builtin(2)
    else:
        V, F, B = M.icosphere(2)
        args = ["screened_poisson", "6", "5", "4"]
        grid = (6, 5, 4)
        eq = Pr.screened_builtin(3, 5.0)
    mesh = tmp_path / "m.obj"
    M.write_obj(str(mesh), V, F, B)
    out = tmp_path / "r.h5"
    cmd = [WOS, *args, str(mesh), "--walks", "32", "--epsilon", "1e-3", "--seed", "4242",
           "-o", str(out)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0
    for line in ("Reading mesh from disk", "Finished BVH build in", "Parallelisation: 1 x 1 x 1",
                 "Beginning walk-on-spheres...", "WoS inside-grid averages:",
                 "Finished walk-on-spheres in", f"Finished writing to {out}."):
        assert line in p.stdout
    if dim == 3:
        assert "3D ray-classification diagnostics:" in p.stdout
    hdr = subprocess.run([H5DUMP, "-H", str(out)], capture_output=True, text=True, timeout=60).stdout
    for ds in ("mean", "variance", "standard_error", "mean_steps", "location", "x", "y", "z"):
        assert f'DATASET "{ds}"' in hdr
    md = ["N_walks", "epsilon", "max_steps", "max_steps_hits", "seed"]
    if dim == 3:
        md += ["max_ray_attempts", "ambiguous_ray_retries", "indeterminate_points", "alpha",
               "source_mode"]
    for name in md:
        assert f'DATASET "{name}"' in hdr
    if dim == 2:
        assert 'DATASET "alpha"' not in hdr  # laplace writes no source metadata
    n = grid[0] * grid[1] * grid[2]
    mean = h5_array(str(out), "/mean", "<f8", str(tmp_path))
    loc = h5_array(str(out), "/location", "u1", str(tmp_path))
    assert mean.size == n and loc.size == n
    ctx = Context(0)
    ctx.set_mesh(V, F, B)
    r = ctx.solve_grid(*grid, eq, Context.params(32, 1e-3, 4242))
    assert np.array_equal(loc, r["location"])
    assert np.array_equal(mean, r["mean"], equal_nan=True)
    x = h5_array(str(out), "/x", "<f8", str(tmp_path))
    assert x[0] == r["summary"]["grid_bounds"][0] and x[-1] == r["summary"]["grid_bounds"][1]
    assert h5_array(str(out), "/metadata/N_walks", "<i4", str(tmp_path))[0] == 32
    assert h5_array(str(out), "/metadata/seed", "<u8", str(tmp_path))[0] == 4242


class ExecOpts(ctypes.Structure):
    """wos_exec_opts of include/wos_run.h."""
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("target_se", ctypes.c_double), ("max_walks", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


def test_registry_entry_point_matches_library(tmp_path):
    """wos_hip_run (libwos_run.so) with the reference's run arguments writes the
    same fields as a direct solve; a problem descriptor and FP32 options pass
    through."""
    C = ctypes
    from wos_amd.capi import WOS_FP32_PERSISTENT
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    lib = C.CDLL(os.path.join(ROOT, "walk-on-spheres_amd", "csrc", "libwos_run.so"))
    lib.wos_hip_run.restype = C.c_int
    V, F, B = M.annulus(64)
    mesh = tmp_path / "a.obj"
    M.write_obj(str(mesh), V, F, B)
    ctx = Context(0)
    ctx.set_mesh(V, F, B)
    for label, eq, opts, prec in (
            ("builtin fp64", None, None, 0),
            ("annulus data fp32", Pr.annulus_laplace(), ExecOpts(WOS_FP32_PERSISTENT, 0, 0.0, 0, 0),
             WOS_FP32_PERSISTENT)):
        out = tmp_path / f"{label.replace(' ', '_')}.h5"
        rc = lib.wos_hip_run(2, str(mesh).encode(), str(out).encode(), 10, 9, 1, 24,
                             C.c_double(1e-3), 1000, 12, C.c_uint64(77), C.c_double(5.0), 0, 0,
                             C.byref(eq) if eq is not None else None,
                             C.byref(opts) if opts is not None else None)
        assert rc == 0, label
        mean = h5_array(str(out), "/mean", "<f8", str(tmp_path))
        want = ctx.solve_grid(10, 9, 1, eq if eq is not None else Pr.laplace_builtin(2),
                              Context.params(24, 1e-3, 77, precision=prec))
        assert np.array_equal(mean, want["mean"], equal_nan=True), label


def test_registry_entry_point_is_reentrant(tmp_path):
    """Two threads call wos_hip_run at once with different execution options
    (FP64 and FP32, different seeds and meshes); each writes what it writes
    when it runs alone (the options are per thread, each call has its own
    context)."""
    import threading
    C = ctypes
    from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    lib = C.CDLL(os.path.join(ROOT, "walk-on-spheres_amd", "csrc", "libwos_run.so"))
    lib.wos_hip_run.restype = C.c_int
    jobs = []
    for k, (mesh_fn, prec, seed) in enumerate(((lambda: M.annulus(64), WOS_FP64_PLAIN, 11),
                                               (lambda: M.star(256, 0.3, 5), WOS_FP32_PERSISTENT, 12))):
        V, F, B = mesh_fn()
        path = tmp_path / f"m{k}.obj"
        M.write_obj(str(path), V, F, B)
        jobs.append((str(path), ExecOpts(prec, 0, 0.0, 0, 0), seed))

    def run(k, tag, rcs):
        path, opts, seed = jobs[k]
        out = str(tmp_path / f"{tag}_{k}.h5")
        rcs[(tag, k)] = lib.wos_hip_run(2, path.encode(), out.encode(), 20, 18, 1, 64,
                                        C.c_double(1e-3), 1000, 12, C.c_uint64(seed),
                                        C.c_double(5.0), 0, 1, None, C.byref(opts))

    rcs = {}
    for k in range(2):
        run(k, "alone", rcs)
    for rep in range(3):
        th = [threading.Thread(target=run, args=(k, f"both{rep}", rcs)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in th)
    assert all(rc == 0 for rc in rcs.values()), rcs
    for k in range(2):
        want = h5_array(str(tmp_path / f"alone_{k}.h5"), "/mean", "<f8", str(tmp_path))
        for rep in range(3):
            got = h5_array(str(tmp_path / f"both{rep}_{k}.h5"), "/mean", "<f8", str(tmp_path))
            assert np.array_equal(got, want, equal_nan=True), (k, rep)


def test_two_contexts_on_two_threads():
    """Two contexts on one GPU, driven from two threads at once: each
    context's solves equal its solves run alone."""
    import threading
    from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
    probs = [(M.icosphere(3), Pr.sphere_screened_constant(), (10, 10, 10), WOS_FP32_PERSISTENT),
             (M.annulus(128), Pr.annulus_laplace(), (40, 40, 1), WOS_FP64_PLAIN)]
    ctxs = []
    for mesh, _, _, _ in probs:
        c = Context(0)
        c.set_mesh(*mesh)
        ctxs.append(c)
    alone = [c.solve_grid(*g, eq, Context.params(128, 1e-3, 5, precision=p))["mean"]
             for c, (_, eq, g, p) in zip(ctxs, probs)]
    got = {}

    def run(k):
        _, eq, g, p = probs[k]
        got[k] = [ctxs[k].solve_grid(*g, eq, Context.params(128, 1e-3, 5, precision=p))["mean"]
                  for _ in range(4)]

    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    for k in range(2):
        assert len(got[k]) == 4
        for m in got[k]:
            assert np.array_equal(m, alone[k], equal_nan=True), k
    for c in ctxs:
        c.close()
