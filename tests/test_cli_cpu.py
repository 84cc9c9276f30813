"""The `wos` command line tool on the CPU: option parsing, registry lookup,
overwrite protection and mesh validation all happen before any GPU call, with
the reference's messages and exit codes (src/main.cpp:22-393,
src/mesh.cpp:51-331)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WOS = os.path.join(ROOT, "walk-on-spheres_amd", "bin", "wos")


def run(*args, cwd=None):
    if not os.path.exists(WOS):
        pytest.fail("wos binary not built; run __graft_entry__.build()")
    p = subprocess.run([WOS, *args], capture_output=True, text=True, cwd=cwd, timeout=60)
    return p.returncode, p.stdout, p.stderr


def test_help_exits_zero():
    rc, out, _ = run("--help")
    assert rc == 0 and out.startswith("Usage:") and "--source-mode MODE" in out


@pytest.mark.parametrize("args,msg", [
    (["--walks", "0"], "Invalid --walks value: 0"),
    (["--walks"], "--walks requires a positive integer"),
    (["--epsilon", "-1"], "Invalid --epsilon value: -1"),
    (["--epsilon", "nan"], "Invalid --epsilon value: nan"),
    (["--max-ray-attempts", "2"], "Invalid --max-ray-attempts value: 2"),
    (["--alpha", "0"], "Invalid --alpha value: 0"),
    (["--source-mode", "mis"], "Invalid --source-mode value: mis"),
    (["--seed", "-3"], "Invalid --seed value: -3"),
    (["--bogus"], "Unknown option: --bogus"),
    (["laplace", "4"], "Nx and Ny must be provided together"),
    (["laplace", "4", "x"], "Nx and Ny must be positive integers"),
    (["laplace", "4", "4", "a", "b"], "Nz must be a positive integer when a fifth positional argument is present"),
    (["laplace", "1", "2", "3", "4", "5", "6"], "Too many positional arguments"),
    (["--precision", "fp16"], "Invalid --precision value: fp16"),
])
def test_option_errors(args, msg):
    rc, _, err = run(*args)
    assert rc == 1 and msg in err


def test_unknown_equation_lists_registry():
    rc, _, err = run("helmholtz3", "4", "4", "m.obj", "--seed", "1")
    assert rc == 1
    assert "Unknown equation: 'helmholtz3'" in err
    for name in ("laplace", "poisson", "screened_poisson", "helmholtz"):
        assert f"  - {name}" in err


def test_refuses_to_overwrite(tmp_path):
    out = tmp_path / "r.h5"
    out.write_text("x")
    rc, _, err = run("laplace", "4", "4", "m.obj", "-o", str(out), "--seed", "1")
    assert rc == 1 and "Refusing to overwrite existing output" in err


def test_warnings_for_ignored_options(tmp_path):
    rc, _, err = run("laplace", "4", "4", str(tmp_path / "none.obj"), "--alpha", "2",
                     "--source-mode", "green", "-o", str(tmp_path / "o.h5"), "--seed", "1")
    assert rc == 1
    assert "--alpha applies only to screened_poisson" in err
    assert "--source-mode applies only to poisson and screened_poisson" in err
    assert "Could not read mesh file" in err


@pytest.mark.parametrize("text,msg", [
    ("v 0 0\nv 1 0\n", "(no 'l' or 'f' entries)"),
    ("v 0 0\nl 1\n", "polyline has fewer than 2 vertices"),
    ("v 0 0\nv 1 0\nl 1 3\n", "vertex index out of range: '3'"),
    ("v 0 0\nv 1 0\nl 1 0\n", "vertex index must be a positive 1-based integer: '0'"),
    ("v 0 0\nv 1 0\nl 1 /2\n", "missing vertex index in '/2'"),
    ("v 0 0\nv 1 0\nv 1 1\nf 1 2 3 3\n", "face has 4 vertices; only triangles are supported"),
    ("v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n", "contains a degenerate triangle"),
    ("v 0 0\nv 0 0\nl 1 2 1\n", "contains a zero-length line segment"),
    ("v 0 x\nv 1 0\nl 1 2\n", "malformed vertex record"),
    ("v 0 0 inf\nv 1 0\nv 1 1\nl 1 2 3\n", "vertex coordinates must be finite"),
])
def test_mesh_validation(tmp_path, text, msg):
    mesh = tmp_path / "m.obj"
    mesh.write_text(text)
    rc, out, err = run("laplace", "4", "4", str(mesh), "-o", str(tmp_path / "o.h5"), "--seed", "1")
    assert rc == 1
    assert msg in err


def test_open_polyline_is_closed_with_warning(tmp_path):
    # the mesh loads (warning), then the run needs a GPU: stop at the context
    mesh = tmp_path / "tri.obj"
    mesh.write_text("v 0 0\nv 1 0\nv 0 1\nl 1 2 3\n")
    rc, out, err = run("laplace", "4", "4", str(mesh), "-o", str(tmp_path / "o.h5"), "--seed", "1")
    assert f"Warning: open polyline in {mesh} auto-closed" in err
    assert "Finished reading." in out
