"""Planar grid cells on the GPU: on a closed box every cell near a face lists
only coplanar triangles, and the FP32 walker's relaxed queries there end at
the plane distance.  An overlarge radius would let a ball cross a face and
bias the source term, so the FP32 estimates are checked against the box's
Poisson solution in closed form (triple sine series) and against the FP64
walker, whose reference traversal has no planar shortcut; the query-unit
counters show that the planar answers are the ones used."""
import math

import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd.capi import (WOS_BND_ZERO, WOS_EQ_POISSON, WOS_FP32_PERSISTENT, WOS_FP64_PLAIN,
                          WOS_SRC_CONSTANT, make_eq)
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

LX, LY, LZ = 1.0, 0.6, 0.4


def box_poisson_exact(x, y, z, terms=61):
    """-Laplace(u) = 1 in the box, u = 0 on its faces:
    u = sum over odd l, m, n of 64 / (pi^3 l m n k^2) sin(l pi x / LX)
    sin(m pi y / LY) sin(n pi z / LZ), k^2 = (l pi / LX)^2 + (m pi / LY)^2 +
    (n pi / LZ)^2 (the product of the 1D sine series of the constant 1)."""
    u = np.zeros_like(x)
    odd = np.arange(1, terms + 1, 2)
    for l in odd:
        sx = np.sin(l * math.pi * x / LX)
        for m in odd:
            sy = np.sin(m * math.pi * y / LY)
            for n in odd:
                k2 = (l * math.pi / LX) ** 2 + (m * math.pi / LY) ** 2 + (n * math.pi / LZ) ** 2
                u += 64.0 / (math.pi ** 3 * l * m * n * k2) * sx * sy * np.sin(n * math.pi * z / LZ)
    return u


@pytest.fixture(scope="module")
def box_ctx():
    ctx = Context(0)
    V, F, _ = M.box(LX, LY, LZ, 30, 18, 12)
    ctx.set_mesh(V, F)
    yield ctx
    ctx.close()


def test_closed_form_pins_the_series():
    # the series at the centre against a much longer one, and zero on a face
    c = box_poisson_exact(np.array([0.5]), np.array([0.3]), np.array([0.2]))
    c_long = box_poisson_exact(np.array([0.5]), np.array([0.3]), np.array([0.2]), terms=161)
    assert abs(c[0] - c_long[0]) < 1e-6
    face = box_poisson_exact(np.array([0.0, 0.4]), np.array([0.3, 0.0]), np.array([0.2, 0.1]))
    assert np.all(np.abs(face) < 1e-12)


def test_planar_cells_match_closed_form_and_fp64(box_ctx):
    eq = make_eq(WOS_EQ_POISSON, source_kind=WOS_SRC_CONSTANT, boundary_kind=WOS_BND_ZERO,
                 src_value=1.0)
    nx, ny, nz = 12, 8, 6
    r32 = box_ctx.solve_grid(nx, ny, nz, eq, Context.params(2048, 1e-3, 17,
                                                             precision=WOS_FP32_PERSISTENT))
    r64 = box_ctx.solve_grid(nx, ny, nz, eq, Context.params(2048, 1e-3, 17,
                                                             precision=WOS_FP64_PLAIN))
    ins = r32["location"] == 1
    assert np.array_equal(r32["location"], r64["location"]) and ins.sum() > 100
    assert r32["summary"]["failed_walks"] == 0
    lo = r32["summary"]["grid_bounds"]
    xs = np.linspace(lo[0], lo[1], nx)
    ys = np.linspace(lo[2], lo[3], ny)
    zs = np.linspace(lo[4], lo[5], nz)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    exact = box_poisson_exact(X.ravel(), Y.ravel(), Z.ravel())[ins]
    # the eps-shell bias: a walk stopped within 1e-3 of a face takes u = 0
    # there, where u is at most about 1e-3 x |grad u| (< 0.2)
    for r in (r32, r64):
        z = (r["mean"][ins] - exact) / np.maximum(r["standard_error"][ins], 1e-12)
        assert np.max(np.abs(r["mean"][ins] - exact) - 4 * r["standard_error"][ins]) < 2e-4
        assert np.mean(z ** 2) < 2.0
    comb = np.sqrt(r32["standard_error"][ins] ** 2 + r64["standard_error"][ins] ** 2)
    z32 = (r32["mean"][ins] - r64["mean"][ins]) / comb
    assert np.max(np.abs(z32)) < 5.0 and np.mean(z32 ** 2) < 2.0


def test_planar_answers_are_used(box_ctx):
    eq = make_eq(WOS_EQ_POISSON, source_kind=WOS_SRC_CONSTANT, boundary_kind=WOS_BND_ZERO,
                 src_value=1.0)
    s = box_ctx.solve_grid(10, 6, 4, eq, Context.params(256, 1e-3, 3, precision=WOS_FP32_PERSISTENT,
                                                       count_units=1))["summary"]
    q = s["q_queries"]
    # nearly every query ends at the cell (a planar cell's plane, or the
    # centre-distance bound); list scans remain where a walk ends (its exact
    # re-query) and near the box's edges
    assert s["q_grid_bound"] > 0.8 * q and s["q_list_scans"] < 0.2 * q
