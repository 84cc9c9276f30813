"""FP32 walker launch shapes and query relaxation.

Every walk draws from its own counter stream and its closest-point queries
depend only on the walker's position, so the FP32 results are bit-identical
for every launch shape (block size, refill chunk, LDS node cache, lanes per
walk-step phase).  A relaxed query (relax > 0) walks with smaller spheres,
which changes the walks but not the expectation: its means agree with the
FP64 plain walker within the combined standard errors.
"""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(block_threads=64, chunk=1, min_ready=0, lds_nodes=0),
    dict(block_threads=128, chunk=4096, min_ready=64, lds_nodes=7),
    dict(block_threads=192, chunk=33, min_ready=1, lds_nodes=-1),
    dict(block_threads=256, chunk=256, min_ready=32, lds_nodes=0, blocks_per_cu=1),
]

CASES = [
    ("icosphere3-screened", lambda: M.icosphere(3), Pr.sphere_screened_constant, (12, 12, 12), 1e-4),
    ("annulus-laplace", M.annulus, Pr.annulus_laplace, (24, 24, 1), 1e-3),
    ("star-gaussian-green", lambda: M.star(512, 0.3, 5), lambda: Pr.star_poisson_gaussian(1),
     (20, 20, 1), 1e-3),
]


@pytest.fixture
def ctx(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_tuning()


@pytest.mark.parametrize("name,mesh_fn,eq_fn,grid,eps", CASES, ids=[c[0] for c in CASES])
def test_fp32_independent_of_launch_shape(ctx, name, mesh_fn, eq_fn, grid, eps):
    V, F, B = mesh_fn()
    ctx.set_mesh(V, F, B)
    eq = eq_fn()
    p = Context.params(64, eps, 5, precision=WOS_FP32_PERSISTENT)
    ctx.set_tuning()
    ref = ctx.solve_grid(*grid, eq, p)
    for shape in SHAPES:
        ctx.set_tuning(**shape)
        got = ctx.solve_grid(*grid, eq, p)
        print(name, shape, got["summary"]["launch_blocks"], got["summary"]["launch_threads"],
              got["summary"]["lds_cached_nodes"])
        for k in ("mean", "variance", "mean_steps", "location", "max_steps_hits"):
            assert np.array_equal(ref[k], got[k], equal_nan=True), (name, shape, k)


@pytest.mark.parametrize("relax", [0.1, 0.3])
@pytest.mark.parametrize("name,mesh_fn,eq_fn,grid,eps", CASES, ids=[c[0] for c in CASES])
def test_fp32_relaxed_query_matches_fp64(ctx, name, mesh_fn, eq_fn, grid, eps, relax):
    V, F, B = mesh_fn()
    ctx.set_mesh(V, F, B)
    eq = eq_fn()
    n = 256
    a = ctx.solve_grid(*grid, eq, Context.params(n, eps, 5, precision=WOS_FP64_PLAIN))
    ctx.set_tuning()
    exact = ctx.solve_grid(*grid, eq, Context.params(n, eps, 5, precision=WOS_FP32_PERSISTENT))
    ctx.set_tuning(relax=relax)
    b = ctx.solve_grid(*grid, eq, Context.params(n, eps, 5, precision=WOS_FP32_PERSISTENT))
    assert np.array_equal(a["location"], b["location"])
    ins = a["location"] == 1
    d = np.abs(a["mean"][ins] - b["mean"][ins])
    comb = np.sqrt(a["standard_error"][ins] ** 2 + b["standard_error"][ins] ** 2) + 1e-12
    z = d / comb
    print(f"{name} relax={relax}: max d/se={z.max():.3f} mean d/se={z.mean():.3f} "
          f"steps {exact['summary']['sum_mean_steps']:.1f} -> {b['summary']['sum_mean_steps']:.1f}")
    # 4 combined SE per point; the mean |z| of a correct estimator is ~0.8
    assert np.all(z <= 4.0), name
    assert z.mean() < 1.2, name
    # smaller spheres take at least as many steps on average
    assert b["summary"]["sum_mean_steps"] >= 0.99 * exact["summary"]["sum_mean_steps"]
    # walks still end in the epsilon shell (no walk stopped by the relaxed bound)
    assert b["summary"]["max_steps_hits"] == exact["summary"]["max_steps_hits"] == 0


def test_tuning_validation(ctx):
    with pytest.raises(RuntimeError):
        ctx.set_tuning(min_ready=65)
    with pytest.raises(RuntimeError):
        ctx.set_tuning(relax=0.5)
    with pytest.raises(RuntimeError):
        ctx.set_tuning(relax=-0.1)
    with pytest.raises(RuntimeError):
        ctx.set_tuning(block_threads=96)
    ctx.set_tuning(min_ready=7, relax=0.25)
    t = ctx.get_tuning()
    assert t["min_ready"] == 7 and abs(t["relax"] - 0.25) < 1e-7
