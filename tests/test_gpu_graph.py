"""Graph replay of a fixed-walk FP32 grid solve (params.use_graph = 1).

The captured sequence (classification, device-side inside-point compaction,
Welford reset, persistent walker, Welford merge) must give the eager solve's
results bit for bit on every replay, with eager work and synchronisations
between the replays, and must follow changes of the problem: another seed
or walk count is captured again; new per-vertex values are read by the
replay itself (the buffer is refilled before each launch).
"""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

FIELDS = ("mean", "variance", "standard_error", "mean_steps", "location", "max_steps_hits",
          "walks", "ambiguous_ray_retries")


def _same(a, b, label):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k], equal_nan=True), (label, k)
    for k in ("inside_count", "total_walks", "total_steps", "max_steps_hits",
              "ambiguous_ray_retries", "indeterminate_points"):
        assert a["summary"][k] == b["summary"][k], (label, k)
    for k in ("sum_variance", "sum_standard_error", "sum_mean_steps"):
        assert a["summary"][k] == b["summary"][k], (label, k)


def _p(n, seed, graph, eps=1e-4):
    return Context.params(n, eps, seed, precision=WOS_FP32_PERSISTENT, use_graph=graph)


def test_graph_replays_equal_eager(gpu_ctx):
    gpu_ctx.set_tuning()
    V, F, _ = M.icosphere(3)
    gpu_ctx.set_mesh(V, F)
    eq = Pr.sphere_screened_constant()
    eager = gpu_ctx.solve_grid(12, 12, 12, eq, _p(128, 7, 0))
    for rep in range(3):
        g = gpu_ctx.solve_grid(12, 12, 12, eq, _p(128, 7, 1))
        _same(eager, g, f"replay {rep}")
        again = gpu_ctx.solve_grid(12, 12, 12, eq, _p(128, 7, 0))
        _same(eager, again, f"eager after replay {rep}")
    # another seed and another walk count: captured again
    for seed, n in ((8, 128), (8, 96)):
        e = gpu_ctx.solve_grid(12, 12, 12, eq, _p(n, seed, 0))
        g = gpu_ctx.solve_grid(12, 12, 12, eq, _p(n, seed, 1))
        _same(e, g, f"seed {seed} n {n}")
        assert not np.array_equal(e["mean"], eager["mean"], equal_nan=True)


def test_graph_follows_vertex_values(gpu_ctx):
    gpu_ctx.set_tuning()
    V, F, _ = M.torus(1.0, 0.35, 40, 20)
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(1)
    for rep in range(3):
        v = rng.normal(size=len(V))
        eq = Pr.laplace_vertex_values(v)
        g = gpu_ctx.solve_grid(16, 16, 8, eq, _p(64, 3, 1, eps=1e-3))
        e = gpu_ctx.solve_grid(16, 16, 8, eq, _p(64, 3, 0, eps=1e-3))
        _same(e, g, f"vertex values {rep}")


def test_graph_batch_equals_eager(gpu_ctx):
    gpu_ctx.set_tuning()
    fam = M.star_family(count=3, seed=5, n_min=128, n_max=1024)
    gpu_ctx.set_domains([f["mesh"] for f in fam])
    eq = Pr.star_poisson_gaussian(1)
    e, te = gpu_ctx.solve_grid_batch(20, 20, 1, eq, _p(48, 21, 0, eps=1e-3))
    for rep in range(2):
        g, tg = gpu_ctx.solve_grid_batch(20, 20, 1, eq, _p(48, 21, 1, eps=1e-3))
        for d in range(3):
            _same(e[d], g[d], f"domain {d} replay {rep}")
        assert te["inside_count"] == tg["inside_count"]


def test_graph_rejections(gpu_ctx):
    V, F, B = M.annulus(64)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.annulus_laplace()
    with pytest.raises(RuntimeError):  # adaptive
        gpu_ctx.solve_grid(8, 8, 1, eq, Context.params(16, 1e-3, 1, precision=WOS_FP32_PERSISTENT,
                                                       target_se=1e-2, max_walks=64, use_graph=1))
    with pytest.raises(RuntimeError):  # FP64
        gpu_ctx.solve_grid(8, 8, 1, eq, Context.params(16, 1e-3, 1, precision=WOS_FP64_PLAIN,
                                                       use_graph=1))
    with pytest.raises(RuntimeError):  # point lists
        gpu_ctx.solve_points(np.array([[0.5, 0.1]]), eq, _p(16, 1, 1, eps=1e-3))
    with pytest.raises(RuntimeError):  # adjoint
        gpu_ctx.adjoint_points(np.array([[0.5, 0.1]]), eq, _p(16, 1, 1, eps=1e-3), np.ones(1))
