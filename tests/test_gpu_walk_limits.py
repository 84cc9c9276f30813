"""Walks end at max_steps.  A walk that stops (max_steps reached, or a radius
below epsilon) after an inexact relaxed step query searches again exactly
before its end; that search may complete inside query_begin (a cell list
whose first entry answers it), and the walk must then end rather than take
another step (max_steps set low makes nearly every walk take this path)."""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN, WOS_FP64_REFERENCE
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

MAX_STEPS = 3


def _check(r, total_walks, total_steps):
    ins = r["location"] == 1
    assert np.all(r["mean_steps"][ins] <= MAX_STEPS)
    assert total_steps <= total_walks * MAX_STEPS
    assert r["max_steps_hits"][ins].sum() > 0


@pytest.mark.timeout(120)
def test_max_steps_ends_walks_2d_batch(gpu_ctx):
    fam = M.star_family(count=4, seed=5, n_min=256, n_max=4096)
    gpu_ctx.set_domains([f["mesh"] for f in fam])
    prm = Context.params(512, 1e-4, 13, max_steps=MAX_STEPS, precision=WOS_FP32_PERSISTENT)
    res, total = gpu_ctx.solve_grid_batch(48, 48, 1, Pr.star_poisson_gaussian(1), prm)
    print("2D batch: walks", total["total_walks"], "steps", total["total_steps"])
    for r in res:
        _check(r, total["total_walks"], total["total_steps"])


@pytest.mark.timeout(120)
@pytest.mark.parametrize("precision", [WOS_FP32_PERSISTENT, WOS_FP64_PLAIN, WOS_FP64_REFERENCE])
def test_max_steps_ends_walks_3d(gpu_ctx, precision):
    # every walker: FP32, FP64 on counter streams, FP64 on the reference's streams
    V, F, _ = M.icosphere(4)
    gpu_ctx.set_mesh(V, F)
    prm = Context.params(512, 1e-4, 17, max_steps=MAX_STEPS, precision=precision)
    r = gpu_ctx.solve_grid(24, 24, 24, Pr.sphere_screened_constant(10.0, 1.0), prm)
    s = r["summary"]
    print("3D: walks", s["total_walks"], "steps", s["total_steps"])
    _check(r, s["total_walks"], s["total_steps"])
