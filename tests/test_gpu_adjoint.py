"""Adjoint replay on the GPU: gradients against finite differences of the
forward solver with common random numbers (same seed: the walks do not
depend on the parameters, so the replayed walks are the forward walks), and
against linearity where the estimator is linear in the data."""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import (WOS_FP32_PERSISTENT, WOS_FP64_PLAIN, WOS_GRAD_ALPHA,
                          WOS_GRAD_BND_VALUES, WOS_GRAD_SRC_CENTER, WOS_GRAD_SRC_RADIUS,
                          WOS_GRAD_SRC_VALUE)
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

PRECISIONS = [WOS_FP64_PLAIN, WOS_FP32_PERSISTENT]


def loss_grad(ctx, pts, eq_fn, prm, target, theta=None):
    """L = 0.5 sum (u - t)^2 over inside points; returns L, dL/du."""
    out = ctx.solve_points(pts, eq_fn(theta), prm)
    u = out["mean"]
    ok = np.isfinite(u)
    r = np.where(ok, u - target, 0.0)
    return 0.5 * float(np.sum(r * r)), r, out


@pytest.mark.parametrize("precision", PRECISIONS)
def test_vertex_gradient_is_exact_linear(gpu_ctx, precision):
    V, F, _ = M.torus(1.0, 0.35, 40, 20)
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(len(V))
    pts = rng.uniform([-1.3, -1.3, -0.3], [1.3, 1.3, 0.3], (300, 3))
    prm = Context.params(64, 1e-3, 3, precision=precision)
    base = gpu_ctx.solve_points(pts, Pr.laplace_vertex_values(v), prm)
    ok = np.isfinite(base["mean"])
    w = np.where(ok, rng.standard_normal(len(pts)), 0.0)
    _, gv, _ = gpu_ctx.adjoint_points(pts, Pr.laplace_vertex_values(v), prm, w)
    for trial in range(3):
        d = rng.standard_normal(len(V))
        up = gpu_ctx.solve_points(pts, Pr.laplace_vertex_values(v + d), prm)["mean"]
        fd = np.sum(w[ok] * (up[ok] - base["mean"][ok]))
        an = float(gv @ d)
        print(f"vertex grad: fd={fd:.12e} adjoint={an:.12e}")
        assert abs(fd - an) <= 1e-9 * max(1.0, abs(fd))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_boundary_id_gradient(gpu_ctx, precision):
    V, F, B = M.annulus(64)
    gpu_ctx.set_mesh(V, F, B)
    pts = np.random.default_rng(1).uniform(-1, 1, (200, 2))

    def eq(g):
        return Pr.make_eq(Pr.WOS_EQ_LAPLACE, boundary_kind=Pr.WOS_BND_ID_CONST, bnd_values=g)

    prm = Context.params(128, 1e-3, 9, precision=precision)
    g0 = (1.0, 0.0)
    u0 = gpu_ctx.solve_points(pts, eq(g0), prm)["mean"]
    ok = np.isfinite(u0)
    w = np.where(ok, 1.0, 0.0)
    _, _, gp = gpu_ctx.adjoint_points(pts, eq(g0), prm, w)
    u1 = gpu_ctx.solve_points(pts, eq((1.0, 1.0)), prm)["mean"]
    assert abs(gp[WOS_GRAD_BND_VALUES + 0] - np.sum(u0[ok])) <= 1e-9 * np.sum(np.abs(u0[ok]))
    assert abs(gp[WOS_GRAD_BND_VALUES + 1] - np.sum(u1[ok] - u0[ok])) <= 1e-9 * len(pts)


def _sphere_eq(theta):
    a, q = (10.0, 1.0) if theta is None else (theta[0], theta[1])
    return Pr.sphere_screened_constant(a, q, 0)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_alpha_and_source_value_gradients(gpu_ctx, precision):
    V, F, _ = M.icosphere(3)
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-0.8, 0.8, (150, 3))
    prm = Context.params(256, 1e-4, 7, precision=precision)
    target = np.zeros(len(pts))
    th = (10.0, 1.0)
    L0, r, out = loss_grad(gpu_ctx, pts, _sphere_eq, prm, target, th)
    _, _, gp = gpu_ctx.adjoint_points(pts, _sphere_eq(th), prm, r)
    # u is linear in the source value: dL/dq = sum r u / q
    ok = np.isfinite(out["mean"])
    assert abs(gp[WOS_GRAD_SRC_VALUE] - np.sum(r[ok] * out["mean"][ok])) <= 1e-9 * abs(gp[WOS_GRAD_SRC_VALUE])
    h = 1e-4
    Lp = loss_grad(gpu_ctx, pts, _sphere_eq, prm, target, (th[0] + h, th[1]))[0]
    Lm = loss_grad(gpu_ctx, pts, _sphere_eq, prm, target, (th[0] - h, th[1]))[0]
    fd = (Lp - Lm) / (2 * h)
    print(f"alpha: fd={fd:.10e} adjoint={gp[WOS_GRAD_ALPHA]:.10e}")
    tol = 1e-6 if precision == WOS_FP64_PLAIN else 2e-3
    assert abs(fd - gp[WOS_GRAD_ALPHA]) <= tol * abs(fd) + 1e-12


@pytest.mark.parametrize("precision", PRECISIONS)
def test_gaussian_source_gradients(gpu_ctx, precision):
    V, F, _ = M.box(2.0, 1.0, 0.5, 12, 6, 4)
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(5)
    pts = rng.uniform([0.1, 0.1, 0.05], [1.9, 0.9, 0.45], (120, 3))
    prm = Context.params(256, 1e-3, 11, precision=precision)
    base = np.array([0.5, 1.0, 0.3, 0.7, 0.5, 0.4])  # alpha, Q, sigma, x0

    def eq(t):
        return Pr.plate_gaussian(alpha=t[0], q=t[1], sigma=t[2], x0=t[3:6])

    target = np.full(len(pts), 0.01)
    L0, r, _ = loss_grad(gpu_ctx, pts, eq, prm, target, base)
    _, _, gp = gpu_ctx.adjoint_points(pts, eq(base), prm, r)
    for idx, k in ((WOS_GRAD_ALPHA, 0), (WOS_GRAD_SRC_RADIUS, 2), (WOS_GRAD_SRC_CENTER, 3),
                   (WOS_GRAD_SRC_CENTER + 1, 4), (WOS_GRAD_SRC_CENTER + 2, 5)):
        h = 1e-4 * max(1.0, abs(base[k]))
        tp, tm = base.copy(), base.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss_grad(gpu_ctx, pts, eq, prm, target, tp)[0] -
              loss_grad(gpu_ctx, pts, eq, prm, target, tm)[0]) / (2 * h)
        print(f"theta[{k}]: fd={fd:.10e} adjoint={gp[idx]:.10e}")
        tol = 1e-5 if precision == WOS_FP64_PLAIN else 5e-3
        assert abs(fd - gp[idx]) <= tol * abs(fd) + 1e-10


def test_alpha_gradient_unavailable_in_green_mode(gpu_ctx):
    V, F, _ = M.icosphere(2)
    gpu_ctx.set_mesh(V, F)
    pts = np.array([[0.1, 0.2, 0.0], [0.0, -0.3, 0.2]])
    _, _, gp = gpu_ctx.adjoint_points(pts, Pr.sphere_screened_constant(10.0, 1.0, 1),
                                      Context.params(16, 1e-3, 1), np.ones(2))
    assert np.isnan(gp[WOS_GRAD_ALPHA]) and np.isfinite(gp[WOS_GRAD_SRC_VALUE])


def test_adjoint_rejects_adaptive(gpu_ctx):
    V, F, B = M.annulus(32)
    gpu_ctx.set_mesh(V, F, B)
    with pytest.raises(RuntimeError, match="fixed number of walks"):
        gpu_ctx.adjoint_points(np.array([[0.5, 0.1]]), Pr.annulus_laplace(),
                               Context.params(16, 1e-3, 1, target_se=1e-2, max_walks=64), np.ones(1))


def test_torch_autograd_matches_adjoint(gpu_ctx):
    torch = pytest.importorskip("torch")
    from wos_amd.autograd import wos_points
    V, F, _ = M.torus(1.0, 0.35, 30, 16)
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(8)
    pts = rng.uniform([-1.3, -1.3, -0.3], [1.3, 1.3, 0.3], (100, 3))
    v = torch.tensor(rng.standard_normal(len(V)), dtype=torch.float64, requires_grad=True)
    prm = Context.params(64, 1e-3, 2, precision=WOS_FP32_PERSISTENT)
    u = wos_points(gpu_ctx, pts, Pr.laplace_vertex_values(np.zeros(len(V))), prm, vertex_values=v)
    ok = torch.isfinite(u)
    loss = 0.5 * torch.sum(torch.where(ok, u, torch.zeros_like(u)) ** 2)
    loss.backward()
    r = np.where(ok.numpy(), u.detach().numpy(), 0.0)
    _, gv, _ = gpu_ctx.adjoint_points(pts, Pr.laplace_vertex_values(v.detach().numpy()), prm, r)
    assert np.allclose(v.grad.numpy(), gv, rtol=1e-12, atol=1e-14)


def test_adaptive_walks_reach_target(gpu_ctx):
    V, F, _ = M.star(1024, 0.3, 5)
    gpu_ctx.set_mesh(V, F)
    eq = Pr.star_poisson_gaussian(1)
    for precision in PRECISIONS:
        prm = Context.params(64, 1e-3, 5, precision=precision, target_se=2e-3, max_walks=8192)
        r = gpu_ctx.solve_grid(24, 24, 1, eq, prm)
        ins = r["location"] == 1
        se, nw = r["standard_error"][ins], r["walks"][ins]
        print(f"adaptive: rounds={r['summary']['rounds']} walks {nw.min()}..{nw.max()} "
              f"se max={se.max():.3e} capped={(nw >= 8192).sum()}")
        assert np.all((se <= 2e-3) | (nw >= 8192))
        assert nw.min() >= 64 and r["summary"]["total_walks"] == nw.sum()
        # a fixed-count run with the same walks reproduces the adaptive statistics
        k = int(np.argmax(nw))
        fixed = gpu_ctx.solve_grid(24, 24, 1, eq, Context.params(int(nw[k]), 1e-3, 5, precision=precision))
        j = np.nonzero(ins)[0][k]
        assert fixed["mean"][j] == r["mean"][j] and fixed["variance"][j] == r["variance"][j]
