"""Batched domains: several meshes in one context, walked in one launch.
Each domain's slice of a batch equals a solve of that domain alone, bit for
bit (same seeds per domain, same arithmetic; only the launch is shared)."""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    for k in ("mean", "variance", "standard_error", "mean_steps", "location", "walks"):
        assert np.array_equal(a[k], b[k], equal_nan=True), (label, k)


@pytest.mark.parametrize("precision", [WOS_FP64_PLAIN, WOS_FP32_PERSISTENT])
def test_batch_equals_separate_2d(gpu_ctx, precision):
    fam = M.star_family(count=5, seed=11, n_min=128, n_max=2048)
    meshes = [f["mesh"] for f in fam]
    eq = Pr.star_poisson_gaussian(1)
    prm = Context.params(48, 1e-3, 21, precision=precision)
    gpu_ctx.set_domains(meshes)
    res, total = gpu_ctx.solve_grid_batch(18, 16, 1, eq, prm)
    assert total["rounds"] == 1
    assert total["inside_count"] == sum(r["summary"]["inside_count"] for r in res)
    for d, (V, F, B) in enumerate(meshes):
        gpu_ctx.set_mesh(V, F, B)
        alone = gpu_ctx.solve_grid(18, 16, 1, eq, prm)
        _same(res[d], alone, f"domain {d}")
        assert res[d]["summary"]["grid_bounds"] == alone["summary"]["grid_bounds"]


@pytest.mark.parametrize("precision", [WOS_FP64_PLAIN, WOS_FP32_PERSISTENT])
def test_batch_equals_separate_3d(gpu_ctx, precision):
    meshes = [M.icosphere(2), M.icosphere(3), M.torus(1.0, 0.35, 30, 14)]
    eq = Pr.sphere_screened_constant(10.0, 1.0, 0)
    prm = Context.params(32, 1e-3, 5, precision=precision)
    gpu_ctx.set_domains(meshes)
    res, _ = gpu_ctx.solve_grid_batch(7, 6, 5, eq, prm)
    for d, (V, F, B) in enumerate(meshes):
        gpu_ctx.set_mesh(V, F, B)
        _same(res[d], gpu_ctx.solve_grid(7, 6, 5, eq, prm), f"domain {d}")


def test_batch_adaptive(gpu_ctx):
    fam = M.star_family(count=4, seed=3, n_min=256, n_max=4096)
    gpu_ctx.set_domains([f["mesh"] for f in fam])
    prm = Context.params(64, 1e-3, 9, precision=WOS_FP32_PERSISTENT, target_se=3e-3, max_walks=4096)
    res, total = gpu_ctx.solve_grid_batch(20, 20, 1, Pr.star_poisson_gaussian(1), prm)
    assert total["rounds"] >= 2
    for r in res:
        ins = r["location"] == 1
        assert np.all((r["standard_error"][ins] <= 3e-3) | (r["walks"][ins] >= 4096))


def test_fp64_adaptive_rounds_equal_fixed(gpu_ctx):
    """FP64 plain over uneven rounds: adaptive rounds split a point's walks,
    and a point that ends with W walks equals a fixed solve with W walks bit
    for bit (a walk's draws depend on its index only; Welford in walk order)."""
    V, F, B = M.icosphere(2)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.sphere_screened_constant(10.0, 1.0, 0)
    ad = gpu_ctx.solve_grid(6, 6, 6, eq, Context.params(16, 1e-3, 3, precision=WOS_FP64_PLAIN,
                                                        target_se=2e-3, max_walks=256))
    assert ad["summary"]["rounds"] >= 2
    ins = ad["location"] == 1
    counts = np.unique(ad["walks"][ins])
    assert len(counts) >= 2
    checked = 0
    for w in counts[:4]:
        fx = gpu_ctx.solve_grid(6, 6, 6, eq, Context.params(int(w), 1e-3, 3, precision=WOS_FP64_PLAIN))
        sel = ins & (ad["walks"] == w)
        for k in ("mean", "variance", "mean_steps"):
            assert np.array_equal(ad[k][sel], fx[k][sel]), (w, k)
        checked += int(sel.sum())
    assert checked > 0
