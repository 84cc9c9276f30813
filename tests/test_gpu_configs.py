"""The benchmark workloads at their full sizes, on the GPU through the C ABI.

Each workload runs at the size it is measured at and is checked against the
oracle on a subsample of its points (FP64 counter streams, same seeds:
per-point statistics agree to rounding) plus properties that hold at full
size: stopping rules, batch = separate solves, linearity of the adjoint,
common-random-number finite differences, and the FP32 walker against FP64.

The reference's own criterion -- the FP64 plain form within 3 combined
standard errors of the reference on the same seeds -- is tested as a
distribution: the GPU's counter streams and the reference's xoshiro streams
(the oracle's reference-stream mode, pinned bit for bit to the compiled
reference's walks) are independent estimates of the same expectation, so
their z-scores must look standard normal.
"""
import numpy as np
import pytest

import pyoracle
from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import (WOS_FP32_PERSISTENT, WOS_FP64_PLAIN, WOS_GRAD_ALPHA,
                          WOS_GRAD_SRC_CENTER, WOS_GRAD_SRC_RADIUS, WOS_GRAD_SRC_VALUE)
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

THREADS = 16
SE = 1e-3


def _grid_pts(summary, grid, dim):
    return pyoracle.grid_points(summary["grid_bounds"], *grid, dim)


def _subsample(ins_idx, n):
    return ins_idx[:: max(1, len(ins_idx) // n)][:n]


def _fp64_vs_oracle(g, o, label, exact_frac=0.97):
    """Counter-stream FP64 on the GPU against the oracle on the same seeds."""
    assert np.array_equal(g["location"], o["location"]), label + ": location codes differ"
    ins = g["location"] == 1
    dm = np.abs(g["mean"][ins] - o["mean"][ins])
    scale = np.maximum(1.0, np.abs(o["mean"][ins]))
    close = dm <= 1e-11 * scale
    se = o["standard_error"][ins]
    print(f"{label}: {ins.sum()} inside, exact-to-1e-11 {close.mean():.4f}, "
          f"max|dmean| {dm.max():.3e}")
    assert close.mean() >= exact_frac, label
    assert np.all(close | (dm <= 3.0 * se)), label


def _z_check(a_mean, a_se, b_mean, b_se, label, max_abs=5.0, z2=(0.0, 1.5), floor=0.0):
    comb = np.sqrt(a_se ** 2 + b_se ** 2)
    d = a_mean - b_mean
    live = comb > floor
    assert np.all(np.abs(d[~live]) <= max(floor, 1e-12)), label
    z = d[live] / comb[live]
    print(f"{label}: n={live.sum()} max|z| {np.abs(z).max():.3f} mean z^2 {np.mean(z ** 2):.3f}")
    assert np.all(np.abs(z) <= max_abs), label
    assert z2[0] <= np.mean(z ** 2) <= z2[1], label
    return z


# ---------------------------------------------------------------------------
# config 1: 2D Laplace annulus -- the "3 combined SE" criterion at full size
# ---------------------------------------------------------------------------
def test_annulus_fp64_within_3se_of_reference_streams(gpu_ctx, oracle):
    V, F, B = M.annulus(128)
    eq = Pr.annulus_laplace()
    gpu_ctx.set_mesh(V, F, B)
    g = gpu_ctx.solve_grid(64, 64, 1, eq, Context.params(1024, 1e-3, 12345, precision=WOS_FP64_PLAIN))
    o = oracle.mesh(V, F, B).solve_grid(eq, 64, 64, 1, 1024, 1e-3, seed=12345, rng_mode=0,
                                        n_threads=THREADS)
    assert np.array_equal(g["location"], o["location"])
    ins = g["location"] == 1
    d = g["mean"][ins] - o["mean"][ins]
    comb = np.sqrt(g["standard_error"][ins] ** 2 + o["standard_error"][ins] ** 2)
    # points within epsilon of the boundary: every walk ends at its start
    assert np.all(d[comb == 0] == 0.0)
    z = d[comb > 0] / comb[comb > 0]
    frac3 = np.mean(np.abs(z) > 3.0)
    print(f"annulus fp64 vs reference streams: {len(z)} points, |z|>3 {frac3:.4f}, "
          f"mean z^2 {np.mean(z ** 2):.3f}, max|z| {np.abs(z).max():.2f}")
    # independent estimates: P(|z| > 3) = 0.27 % for a standard normal
    assert frac3 <= 0.01
    assert 0.85 <= np.mean(z ** 2) <= 1.15
    assert np.abs(z).max() <= 5.0


# ---------------------------------------------------------------------------
# config 2: icosphere s4, 32^3 grid, 4096 walks
# ---------------------------------------------------------------------------
def test_icosphere_s4_subsample_matches_oracle(gpu_ctx, oracle):
    V, F, _ = M.icosphere(4)
    eq = Pr.sphere_screened_constant(alpha=10.0, f=1.0)
    grid = (32, 32, 32)
    gpu_ctx.set_mesh(V, F)
    r32 = gpu_ctx.solve_grid(*grid, eq, Context.params(4096, 1e-4, 7, precision=WOS_FP32_PERSISTENT))
    ins_idx = np.nonzero(r32["location"] == 1)[0]
    assert len(ins_idx) == 15408
    pick = _subsample(ins_idx, 256)
    pts = _grid_pts(r32["summary"], grid, 3)[pick]
    g = gpu_ctx.solve_points(pts, eq, Context.params(4096, 1e-4, 7, precision=WOS_FP64_PLAIN),
                             seed_index=pick)
    o = oracle.mesh(V, F).solve_points(eq, pts, 4096, 1e-4, seed=7, rng_mode=1, n_threads=THREADS,
                                       seed_index=pick)
    _fp64_vs_oracle(g, o, "icosphere s4 fp64")
    # the FP32 grid walks on the same seeds (common random numbers)
    _z_check(r32["mean"][pick], r32["standard_error"][pick], g["mean"], g["standard_error"],
             "icosphere s4 fp32 grid vs fp64", max_abs=3.0, z2=(0.0, 0.5))
    # the reference's own streams: an independent estimate of the same field
    ref = oracle.mesh(V, F).solve_points(eq, pts, 4096, 1e-4, seed=7, rng_mode=0,
                                         n_threads=THREADS, seed_index=pick)
    _z_check(g["mean"], g["standard_error"], ref["mean"], ref["standard_error"],
             "icosphere s4 fp64 vs reference streams", max_abs=5.0, z2=(0.5, 1.6))


def test_indexed_points_equal_grid_points(gpu_ctx):
    """wos_solve_points_indexed gives a grid point its grid seed: a subset of
    the grid solved as a point list equals the grid solve there bit for bit."""
    V, F, _ = M.icosphere(4)
    eq = Pr.sphere_screened_constant(alpha=10.0, f=1.0)
    gpu_ctx.set_mesh(V, F)
    for prec in (WOS_FP64_PLAIN, WOS_FP32_PERSISTENT):
        prm = Context.params(64, 1e-4, 7, precision=prec)
        full = gpu_ctx.solve_grid(32, 32, 32, eq, prm)
        pick = np.random.default_rng(0).choice(32 ** 3, 500, replace=False)
        pts = _grid_pts(full["summary"], (32, 32, 32), 3)[pick]
        sub = gpu_ctx.solve_points(pts, eq, prm, seed_index=pick)
        for k in ("mean", "variance", "mean_steps", "location", "ambiguous_ray_retries"):
            assert np.array_equal(sub[k], full[k][pick], equal_nan=True), (prec, k)


# ---------------------------------------------------------------------------
# config 3: 16 batched star domains, 256^2 each, adaptive 64 -> 16384 walks
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def stars_run():
    fam = M.star_family(16, seed=2024, n_min=512, n_max=16384)
    ctx = Context(0)
    ctx.set_domains([f["mesh"] for f in fam])
    eq = Pr.star_poisson_gaussian(1)
    prm = Context.params(64, 1e-3, 11, precision=WOS_FP32_PERSISTENT, target_se=SE,
                         max_walks=16384)
    res, tot = ctx.solve_grid_batch(256, 256, 1, eq, prm)
    yield fam, eq, prm, res, tot, ctx
    ctx.close()


def test_stars_adaptive_stopping_rule(stars_run):
    fam, _, _, res, tot, _ = stars_run
    print(f"stars: rounds {tot['rounds']}, walks {tot['total_walks']}, steps {tot['total_steps']}")
    assert 2 <= tot["rounds"] <= 64
    assert tot["failed_walks"] == 0
    inside = 0
    for f, r in zip(fam, res):
        ins = r["location"] == 1
        inside += ins.sum()
        se, nw = r["standard_error"][ins], r["walks"][ins]
        assert np.all((se <= SE) | (nw >= 16384)), f["n"]
        assert np.all(nw >= 64) and np.all(nw <= 16384)
        assert np.all(np.isfinite(r["mean"][ins]))
    assert tot["inside_count"] == inside and tot["total_walks"] == sum(
        int(r["walks"][r["location"] == 1].sum()) for r in res)


def test_stars_batch_slices_equal_separate_solves(stars_run):
    fam, eq, prm, res, _, ctx = stars_run
    order = np.argsort([f["n"] for f in fam])
    alone = Context(0)
    for d in (int(order[0]), int(order[-1])):   # the smallest and the largest domain
        alone.set_mesh(*fam[d]["mesh"])
        r = alone.solve_grid(256, 256, 1, eq, prm)
        for k in ("mean", "variance", "standard_error", "mean_steps", "location", "walks"):
            assert np.array_equal(r[k], res[d][k], equal_nan=True), (d, k)
    alone.close()


def test_stars_fp64_subsample_matches_oracle(stars_run, oracle):
    fam, eq, _, res, _, _ = stars_run
    gpu = Context(0)
    for d in range(16):
        V, F, B = fam[d]["mesh"]
        ins_idx = np.nonzero(res[d]["location"] == 1)[0]
        pick = _subsample(ins_idx, 64)
        pts = _grid_pts(res[d]["summary"], (256, 256, 1), 2)[pick]
        gpu.set_mesh(V, F, B)
        g = gpu.solve_points(pts, eq, Context.params(1024, 1e-3, 11, precision=WOS_FP64_PLAIN),
                             seed_index=pick)
        o = oracle.mesh(V, F, B).solve_points(eq, pts, 1024, 1e-3, seed=11, rng_mode=1,
                                              n_threads=THREADS, seed_index=pick)
        _fp64_vs_oracle(g, o, f"star {d} ({fam[d]['n']} segments) fp64")
        # the adaptive FP32 batch estimate of the same points (its first 1024
        # walks are the FP64 walks' FP32 twins; the rest are new)
        _z_check(res[d]["mean"][pick], res[d]["standard_error"][pick], g["mean"],
                 g["standard_error"], f"star {d} fp32 adaptive vs fp64", max_abs=4.0,
                 z2=(0.0, 1.5))
    gpu.close()


# ---------------------------------------------------------------------------
# config 4: torus, 40 000 triangles, 10 000 probes, 2048 walks
# ---------------------------------------------------------------------------
def _torus_probes(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(o) for o in out) < n:
        p = rng.uniform([-1.35, -1.35, -0.35], [1.35, 1.35, 0.35], size=(4 * n, 3))
        q = (np.hypot(p[:, 0], p[:, 1]) - 1.0) ** 2 + p[:, 2] ** 2 < 0.35 ** 2
        out.append(p[q])
    return np.concatenate(out)[:n]


@pytest.mark.parametrize("precision", [WOS_FP64_PLAIN, WOS_FP32_PERSISTENT])
def test_torus_full_adjoint_is_linear(gpu_ctx, precision):
    V, F, _ = M.torus(1.0, 0.35, 200, 100)
    assert len(F) == 40000
    gpu_ctx.set_mesh(V, F)
    rng = np.random.default_rng(3)
    v = rng.normal(size=len(V))
    vt = v + 0.1 * rng.normal(size=len(V))
    pts = _torus_probes(10000, 3)
    prm = Context.params(2048, 1e-3, 3, precision=precision)
    target = gpu_ctx.solve_points(pts, Pr.laplace_vertex_values(vt), prm)["mean"]
    base = gpu_ctx.solve_points(pts, Pr.laplace_vertex_values(v), prm)
    ok = np.isfinite(base["mean"])
    # the probes are uniform in the smooth torus; the few between it and the
    # inscribed facets are outside the mesh (NaN)
    assert ok.sum() >= 9950
    dl = np.where(ok, base["mean"] - target, 0.0)
    out, gv, _ = gpu_ctx.adjoint_points(pts, Pr.laplace_vertex_values(v), prm, dl)
    assert np.array_equal(out["mean"], base["mean"], equal_nan=True)  # the replay is the forward solve
    assert out["summary"]["failed_walks"] == 0
    for trial in range(3):
        d = rng.standard_normal(len(V))
        up = gpu_ctx.solve_points(pts, Pr.laplace_vertex_values(v + d), prm)["mean"]
        fd = float(np.sum(dl[ok] * (up[ok] - base["mean"][ok])))
        an = float(gv @ d)
        print(f"torus vertex grad: fd={fd:.12e} adjoint={an:.12e}")
        assert abs(fd - an) <= 1e-9 * max(1.0, abs(fd))


def test_torus_fp64_probes_match_oracle(gpu_ctx, oracle):
    V, F, _ = M.torus(1.0, 0.35, 200, 100)
    v = np.random.default_rng(3).normal(size=len(V))
    eq = Pr.laplace_vertex_values(v)
    pts = _torus_probes(10000, 3)[:200]
    gpu_ctx.set_mesh(V, F)
    g = gpu_ctx.solve_points(pts, eq, Context.params(2048, 1e-3, 3, precision=WOS_FP64_PLAIN))
    o = oracle.mesh(V, F).solve_points(eq, pts, 2048, 1e-3, seed=3, rng_mode=1, n_threads=THREADS)
    _fp64_vs_oracle(g, o, "torus 40k fp64 probes")


# ---------------------------------------------------------------------------
# config 5: the 200k-triangle plate, 64 x 32 x 16, 8192 walks
# ---------------------------------------------------------------------------
PLATE_GRID = (64, 32, 16)


def _plate_eq(t):
    return Pr.plate_gaussian(alpha=t[0], q=t[1], sigma=t[2], x0=tuple(t[3:6]))


@pytest.fixture(scope="module")
def plate_ctx():
    V, F, _ = M.plate_200k()
    ctx = Context(0)
    ctx.set_mesh(V, F)
    yield (V, F), ctx
    ctx.close()


def test_plate_fp64_subsample_matches_oracle_and_fp32(plate_ctx, oracle):
    (V, F), ctx = plate_ctx
    eq = _plate_eq([0.05, 1.0, 3.0, 40.0, 30.0, 10.0])
    r32 = ctx.solve_grid(*PLATE_GRID, eq, Context.params(8192, 0.01, 5, precision=WOS_FP32_PERSISTENT))
    assert r32["summary"]["failed_walks"] == 0
    ins_idx = np.nonzero(r32["location"] == 1)[0]
    # points nearest the source carry the signal (far ones are ~1e-100)
    pts_all = _grid_pts(r32["summary"], PLATE_GRID, 3)
    near = ins_idx[np.argsort(np.linalg.norm(pts_all[ins_idx] - [40.0, 30.0, 10.0], axis=1))]
    pick = np.sort(np.concatenate([near[:150], _subsample(ins_idx, 50)]))
    pick = np.unique(pick)
    pts = pts_all[pick]
    g = ctx.solve_points(pts, eq, Context.params(8192, 0.01, 5, precision=WOS_FP64_PLAIN),
                         seed_index=pick)
    o = oracle.mesh(V, F).solve_points(eq, pts, 8192, 0.01, seed=5, rng_mode=1, n_threads=THREADS,
                                       seed_index=pick)
    _fp64_vs_oracle(g, o, "plate 200k fp64")
    floor = 1e-9 * np.abs(g["mean"]).max()
    _z_check(r32["mean"][pick], r32["standard_error"][pick], g["mean"], g["standard_error"],
             "plate fp32 grid vs fp64", max_abs=4.0, z2=(0.0, 1.0), floor=floor)


def test_plate_full_gradients(plate_ctx):
    _, ctx = plate_ctx
    base = np.array([0.05, 1.0, 3.0, 40.0, 30.0, 10.0])  # alpha, Q, sigma, x0
    prm = Context.params(8192, 0.01, 5, precision=WOS_FP32_PERSISTENT)
    r = ctx.solve_grid(*PLATE_GRID, _plate_eq(base), prm)
    ins = r["location"] == 1
    u = r["mean"]
    dl = np.where(ins, 1.0, 0.0)       # L = sum of u over the inside points
    _, _, gp = ctx.adjoint_grid(*PLATE_GRID, _plate_eq(base), prm, dl)
    # u is linear in Q: dL/dQ = sum u / Q exactly (the same walks)
    dq = float(np.sum(u[ins])) / base[1]
    print(f"plate dL/dQ adjoint {gp[WOS_GRAD_SRC_VALUE]:.12e} sum u/Q {dq:.12e}")
    assert abs(gp[WOS_GRAD_SRC_VALUE] - dq) <= 1e-9 * abs(dq)

    def L(t):
        rr = ctx.solve_grid(*PLATE_GRID, _plate_eq(t), prm)
        return float(np.sum(rr["mean"][ins]))

    for idx, k, h in ((WOS_GRAD_SRC_RADIUS, 2, 1e-2), (WOS_GRAD_SRC_CENTER, 3, 1e-2),
                      (WOS_GRAD_SRC_CENTER + 1, 4, 1e-2), (WOS_GRAD_SRC_CENTER + 2, 5, 1e-2),
                      (WOS_GRAD_ALPHA, 0, 1e-4)):
        tp, tm = base.copy(), base.copy()
        tp[k] += h
        tm[k] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        print(f"plate theta[{k}]: fd={fd:.8e} adjoint={gp[idx]:.8e}")
        # FP32 walks; central differences with common random numbers
        assert abs(fd - gp[idx]) <= 5e-3 * abs(fd) + 1e-6 * abs(dq), k
