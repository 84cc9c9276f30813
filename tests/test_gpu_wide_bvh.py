"""The FP32 walker's 4-wide tree (bvh_width = 4) and the LDS share of its
traversal stack (stack_lds).

The 4-wide tree folds every second level of the binary tree into one node,
so its boxes are boxes of the binary tree and an exact query (relax = 0)
finds the same nearest distance up to the ulp-level ties between a box and a
primitive inside it.  A relaxed query stops once its lower bound reaches
(1 - relax) x the best distance found; which primitive that is depends on the
visiting order, so relaxed walks differ between the two trees but keep the
expectation.  Where a stack entry lives (LDS or global memory) changes
nothing at all: the results are bit-identical for every stack_lds.
"""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

MESHES = [
    ("icosphere3", lambda: M.icosphere(3)),
    ("torus", lambda: M.torus(1.0, 0.35, 40, 20)),
    ("annulus", lambda: M.annulus()),
    ("star2048", lambda: M.star(2048, 0.3, 7)),
]

WALKS = [
    ("icosphere3-screened", lambda: M.icosphere(3), Pr.sphere_screened_constant, (12, 12, 12), 1e-4),
    ("annulus-laplace", M.annulus, Pr.annulus_laplace, (24, 24, 1), 1e-3),
    ("star-gaussian-green", lambda: M.star(512, 0.3, 5), lambda: Pr.star_poisson_gaussian(1),
     (20, 20, 1), 1e-3),
]


@pytest.fixture
def ctx(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_tuning()


def _points(V, dim, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = V[:, :dim].min(0), V[:, :dim].max(0)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * 1.3
    p = c + h * (2 * rng.random((n, dim)) - 1)
    k = min(len(V), n // 8)
    p[:k] = V[:k, :dim]
    p[k:2 * k] = V[:k, :dim] * (1 + 1e-7)
    return p


@pytest.mark.parametrize("name,mesh_fn", MESHES, ids=[c[0] for c in MESHES])
def test_wide_tree_closest_points(ctx, name, mesh_fn):
    V, F, B = mesh_fn()
    dim = F.shape[1]
    ctx.set_mesh(V, F, B)
    pts = _points(np.asarray(V), dim, 3000, 23)
    d64 = ctx.closest_points(pts, WOS_FP64_PLAIN)["dist"]
    margin = 32 * 2.0 ** -24 * max(1.0, np.abs(V).max())
    for grid_cells in (0, 1 << 16):
        ctx.set_tuning(grid_cells=grid_cells, bvh_width=2)
        r2 = ctx.closest_points(pts, WOS_FP32_PERSISTENT)
        ctx.set_tuning(grid_cells=grid_cells, bvh_width=4)
        r4 = ctx.closest_points(pts, WOS_FP32_PERSISTENT)
        same = np.mean(r4["dist"] == r2["dist"])
        print(f"{name} grid {grid_cells}: width 4 == width 2 for {same:.5f} of the points, "
              f"max|d4 - d64| {np.abs(r4['dist'] - d64).max():.3e} (margin {margin:.3e})")
        assert np.all(np.abs(r4["dist"] - d64) <= margin), (name, grid_cells)
        assert np.array_equal(r4["lower"], r4["dist"]), (name, grid_cells)
        assert same >= 0.99, (name, grid_cells)
        dproj = np.linalg.norm(r4["nearest"] - pts, axis=1)
        assert np.all(np.abs(dproj - d64) <= margin), (name, grid_cells)
        for relax in (0.05, 0.2, 0.45):
            r = ctx.closest_points(pts, WOS_FP32_PERSISTENT, relax=relax)
            assert np.all(r["lower"] <= d64 + margin), (name, grid_cells, relax)
            assert np.all(r["lower"] >= (1 - relax) * r["dist"] * (1 - 1e-6)), (name, grid_cells, relax)
            assert np.all(r["dist"] >= d64 - margin), (name, grid_cells, relax)
            found = r["prim"] >= 0
            assert np.all(np.isnan(r["nearest"][~found])), (name, grid_cells, relax)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("name,mesh_fn,eq_fn,grid,eps", WALKS, ids=[c[0] for c in WALKS])
def test_stack_lds_is_result_neutral(ctx, name, mesh_fn, eq_fn, grid, eps, width):
    V, F, B = mesh_fn()
    ctx.set_mesh(V, F, B)
    eq = eq_fn()
    p = Context.params(64, eps, 5, precision=WOS_FP32_PERSISTENT)
    ref = None
    for grid_cells in (0, -1):
        for stack_lds in (0, 1, 8, 32, 64):
            ctx.set_tuning(bvh_width=width, stack_lds=stack_lds, grid_cells=grid_cells)
            got = ctx.solve_grid(*grid, eq, p)
            if stack_lds == 0:
                ref = got
                continue
            for k in ("location", "mean", "variance", "mean_steps"):
                assert np.array_equal(ref[k], got[k], equal_nan=True), (name, grid_cells, stack_lds, k)


@pytest.mark.parametrize("name,mesh_fn,eq_fn,grid,eps", WALKS, ids=[c[0] for c in WALKS])
def test_wide_tree_walks(ctx, name, mesh_fn, eq_fn, grid, eps):
    V, F, B = mesh_fn()
    ctx.set_mesh(V, F, B)
    eq = eq_fn()
    # exact tree queries: the two trees give the same spheres up to ulp ties
    p = Context.params(64, eps, 5, precision=WOS_FP32_PERSISTENT)
    ctx.set_tuning(grid_cells=0, relax=0.0, bvh_width=2)
    a = ctx.solve_grid(*grid, eq, p)
    ctx.set_tuning(grid_cells=0, relax=0.0, bvh_width=4)
    b = ctx.solve_grid(*grid, eq, p)
    assert np.array_equal(a["location"], b["location"])
    fin = ~np.isnan(a["mean"])
    for k in ("mean", "variance", "mean_steps"):
        same = np.mean(a[k][fin] == b[k][fin])
        print(name, k, "bit-identical fraction", same)
        assert same >= 0.95, (name, k)
    # relaxed queries (the default shape): the expectation of the FP64 walker
    n = 256
    ref = ctx.solve_grid(*grid, eq, Context.params(n, eps, 5, precision=WOS_FP64_PLAIN))
    for grid_cells in (0, -1):
        ctx.set_tuning(bvh_width=4, grid_cells=grid_cells)
        r = ctx.solve_grid(*grid, eq, Context.params(n, eps, 5, precision=WOS_FP32_PERSISTENT))
        assert np.array_equal(ref["location"], r["location"])
        ins = ref["location"] == 1
        d = np.abs(ref["mean"][ins] - r["mean"][ins])
        comb = np.sqrt(ref["standard_error"][ins] ** 2 + r["standard_error"][ins] ** 2) + 1e-12
        z = d / comb
        print(f"{name} width 4 grid {grid_cells}: max d/se={z.max():.3f} mean d/se={z.mean():.3f}")
        assert np.all(z <= 4.0), (name, grid_cells)
        assert z.mean() < 1.2, (name, grid_cells)
        assert r["summary"]["max_steps_hits"] == 0
        assert r["summary"]["failed_walks"] == 0


def test_wide_tree_tuning_validation(ctx):
    for bad in (dict(bvh_width=3), dict(bvh_width=8), dict(stack_lds=-1), dict(stack_lds=65)):
        with pytest.raises(RuntimeError):
            ctx.set_tuning(**bad)
    ctx.set_tuning(bvh_width=4, stack_lds=16)
    t = ctx.get_tuning()
    assert t["bvh_width"] == 4 and t["stack_lds"] == 16


def test_autotune_searches_the_tree_width(ctx):
    V, F, B = M.icosphere(3)
    ctx.set_mesh(V, F, B)
    eq = Pr.sphere_screened_constant()
    grid = (16, 16, 16)
    p = Context.params(64, 1e-4, 5, precision=WOS_FP32_PERSISTENT, tune=2)
    ctx.set_tuning()
    default = ctx.get_tuning()
    chosen = ctx.autotune(*grid, eq, p)
    print("autotuned (tune=2)", chosen)
    assert chosen["bvh_width"] in (2, 4)
    assert chosen["stack_lds"] in (0, 8, 32)
    # a tune=2 solve runs the chosen shape and leaves the context's own
    tuned = ctx.solve_grid(*grid, eq, p)
    after = ctx.get_tuning()
    ctx.set_tuning(**chosen)
    direct = ctx.solve_grid(*grid, eq, Context.params(64, 1e-4, 5, precision=WOS_FP32_PERSISTENT))
    for k in ("location", "mean", "variance", "mean_steps"):
        assert np.array_equal(tuned[k], direct[k], equal_nan=True), k
    assert after == chosen  # autotune() installs its choice
    ctx.set_tuning()
    assert ctx.get_tuning() == default
