"""Green-source MIS (source mode WOS_SOURCE_GREEN_MIS) through the C ABI.

The welding level-1 problem (apps/welding/src/level1_stationary_heat.cpp:
42-62, :624-666): 2D Poisson on the plate [0, 0.12] x [0, 0.08], g = 0, a
finite-plate Gaussian source, solved in the Green mode and with the
balance-heuristic mixture of include/wos/source_sampling.hpp:35-78.
* FP64 plain against the oracle's counter streams: per-point statistics to
  rounding (OCML and glibc transcendentals differ in the last bit);
* the reference's own walks: tests/golden/walks.npz holds them for four
  welding / MIS cases, checked in the reference-stream tests;
* FP32 persistent against FP64 plain, and MIS against the Green mode: the
  same field within combined standard errors, MIS with the lower variance;
* descriptor checks: only 2D Gaussian sources of poisson / screened_poisson,
  beta in (0, 1), a box around the source centre, no adjoint.
"""
import math

import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import (WOS_FP32_PERSISTENT, WOS_FP64_PLAIN, WOS_SOURCE_GREEN,
                          WOS_SOURCE_GREEN_MIS, WOS_SRC_CONSTANT)
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

PLATE = M.rectangle(0.0, 0.12, 0.0, 0.08)
EPS = 8e-6


@pytest.mark.parametrize("case", ["box", "plane", "screened"])
def test_mis_fp64_matches_oracle(gpu_ctx, oracle, case):
    eq = {"box": lambda: Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS),
          "plane": lambda: Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS, 0.3, truncate=False),
          "screened": lambda: Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS, 0.7, alpha=40.0)}[case]()
    gpu_ctx.set_mesh(*PLATE)
    grid = (25, 17, 1)
    g = gpu_ctx.solve_grid(*grid, eq, Context.params(128, EPS, 2026, precision=WOS_FP64_PLAIN))
    o = oracle.mesh(*PLATE).solve_grid(eq, *grid, 128, EPS, seed=2026, rng_mode=1, n_threads=16)
    assert np.array_equal(g["location"], o["location"])
    ins = g["location"] == 1
    assert ins.sum() >= 300
    dm = np.abs(g["mean"][ins] - o["mean"][ins])
    scale = np.maximum(1.0, np.abs(o["mean"][ins]))
    close = dm <= 1e-11 * scale
    se = np.maximum(o["standard_error"][ins], 1e-300)
    print(f"mis {case}: exact-to-1e-11 {close.mean():.4f} max|dmean| {dm.max():.3e} "
          f"max dmean/se {np.max(dm / se):.3f}")
    assert np.all(np.isfinite(g["mean"][ins]))
    assert close.mean() >= 0.97
    assert np.all(dm <= 3.0 * se + 1e-11 * scale)
    assert np.mean(g["mean_steps"][ins] == o["mean_steps"][ins]) >= 0.97
    b = g["location"] == 2
    assert np.all(g["mean"][b] == 0.0)


def test_mis_fp32_and_green_estimate_the_same_field(gpu_ctx):
    """Full level-1 grid (49 x 33) at 2048 walks: FP32 MIS against FP64 MIS,
    and MIS against the Green mode, within combined standard errors; MIS
    has the lower variance (the reason for it)."""
    gpu_ctx.set_mesh(*PLATE)
    cfg = Pr.WELDING_L1
    grid = (cfg["nx"], cfg["ny"], 1)
    n = 2048
    mis = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    a = gpu_ctx.solve_grid(*grid, mis, Context.params(n, EPS, 7, precision=WOS_FP64_PLAIN))
    b = gpu_ctx.solve_grid(*grid, mis, Context.params(n, EPS, 7, precision=WOS_FP32_PERSISTENT))
    c = gpu_ctx.solve_grid(*grid, Pr.welding_gaussian(WOS_SOURCE_GREEN),
                           Context.params(n, EPS, 7, precision=WOS_FP32_PERSISTENT))
    assert np.array_equal(a["location"], b["location"])
    ins = a["location"] == 1
    for x, label in ((b, "fp32 mis vs fp64 mis"), (c, "fp32 green vs fp64 mis")):
        comb = np.sqrt(a["standard_error"][ins] ** 2 + x["standard_error"][ins] ** 2)
        z = (a["mean"][ins] - x["mean"][ins]) / comb
        print(f"{label}: max|z| {np.abs(z).max():.3f} mean z^2 {np.mean(z ** 2):.3f}")
        assert np.all(np.abs(z) <= 5.0), label
        assert np.mean(z ** 2) < 1.4, label
    ratio = b["variance"][ins] / c["variance"][ins]
    print(f"variance MIS / Green: median {np.median(ratio):.3f} max {ratio.max():.3f}")
    assert np.median(ratio) < 0.8
    # the field is symmetric about the plate centre
    u = b["mean"].reshape(grid[0], grid[1])
    flip = u[::-1, ::-1]
    se = b["standard_error"].reshape(u.shape)
    i = ins.reshape(u.shape)
    d = np.abs(u - flip)[i] / np.maximum(np.hypot(se, se[::-1, ::-1])[i], 1e-300)
    assert np.all(d <= 5.0)


def test_mis_reproducible_and_point_lists(gpu_ctx):
    gpu_ctx.set_mesh(*PLATE)
    eq = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    p = Context.params(256, EPS, 11, precision=WOS_FP32_PERSISTENT)
    r1 = gpu_ctx.solve_grid(13, 9, 1, eq, p)
    r2 = gpu_ctx.solve_grid(13, 9, 1, eq, p)
    assert np.array_equal(r1["mean"], r2["mean"], equal_nan=True)
    # the grid's own coordinates, lo + (hi - lo) i / (n - 1)
    gb = r1["summary"]["grid_bounds"]
    pts = np.array([(gb[0] + (gb[1] - gb[0]) * i / 12, gb[2] + (gb[3] - gb[2]) * j / 8)
                    for i in range(13) for j in range(9)])
    r3 = gpu_ctx.solve_points(pts, eq, p)
    assert np.array_equal(r1["mean"], r3["mean"], equal_nan=True)


def test_mis_descriptor_checks(gpu_ctx):
    gpu_ctx.set_mesh(*PLATE)
    p = Context.params(8, EPS, 1)
    pts = np.array([[0.06, 0.04]])

    def refused(eq, prm=p):
        with pytest.raises(RuntimeError):
            gpu_ctx.solve_points(pts, eq, prm)

    for beta in (0.0, 1.0, -0.5, 1.5, math.nan, math.inf):
        refused(Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS, beta))
    bad = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    bad.proposal_box[0], bad.proposal_box[1] = 0.1, 0.05    # inverted
    refused(bad)
    bad = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    bad.proposal_box[1] = 0.03                              # centre outside
    refused(bad)
    bad = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    bad.source_kind = WOS_SRC_CONSTANT                      # no proposal
    refused(bad)
    bad = Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS)
    bad.source_mode = 3
    refused(bad)
    for prec in (WOS_FP64_PLAIN, WOS_FP32_PERSISTENT):
        with pytest.raises(RuntimeError):
            gpu_ctx.adjoint_points(pts, Pr.welding_gaussian(WOS_SOURCE_GREEN_MIS),
                                   Context.params(8, EPS, 1, precision=prec), np.ones(1))
    # 3D sources have no proposal
    V, F, B = M.icosphere(1)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.sphere_screened_constant(10.0, 1.0, WOS_SOURCE_GREEN_MIS)
    with pytest.raises(RuntimeError):
        gpu_ctx.solve_points(np.zeros((1, 3)), eq, p)
    # without a source the mode is unused (as for Laplace in the reference)
    gpu_ctx.set_mesh(*PLATE)
    eq = Pr.laplace_builtin(2)
    eq.source_mode = WOS_SOURCE_GREEN_MIS
    r = gpu_ctx.solve_points(pts, eq, p)
    assert np.isfinite(r["mean"][0])
