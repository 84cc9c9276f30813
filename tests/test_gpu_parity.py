"""GPU parity: the HIP library through its C ABI against the C oracle.

The FP64 plain walker uses the same counter streams as the oracle's counter
mode and restates its arithmetic, so per-point statistics agree to rounding
(a walk can only differ where OCML and glibc transcendentals differ by an
ulp and flip a termination test).  Location codes are compared bit-exactly.
The FP32 persistent walker draws the same uniforms (top 24 bits), so its
per-point means track the FP64 ones closely (common random numbers).
"""
import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

THREADS = 16


def _run_both(ctx, om, mesh, eq, grid, n_walks, eps, seed, precision=WOS_FP64_PLAIN,
              max_steps=1000):
    V, F, B = mesh
    ctx.set_mesh(V, F, B)
    g = ctx.solve_grid(*grid, eq, Context.params(n_walks, eps, seed, max_steps=max_steps,
                                                 precision=precision))
    o = om.solve_grid(eq, *grid, n_walks, eps, max_steps=max_steps, seed=seed, rng_mode=1,
                      n_threads=THREADS)
    return g, o


def _compare_fp64(g, o, label, exact_frac=0.97):
    assert np.array_equal(g["location"], o["location"]), label + ": location codes differ"
    ins = g["location"] == 1
    assert ins.any()
    dm = np.abs(g["mean"][ins] - o["mean"][ins])
    scale = np.maximum(1.0, np.abs(o["mean"][ins]))
    close = dm <= 1e-11 * scale
    se = np.maximum(o["standard_error"][ins], 1e-300)
    print(f"{label}: inside={ins.sum()} exact-to-1e-11={close.mean():.4f} "
          f"max|dmean|={dm.max():.3e} max dmean/se={np.max(dm / se):.3f}")
    assert close.mean() >= exact_frac, label
    assert np.all(dm <= 3.0 * se + 1e-11 * scale), label
    b = g["location"] == 2
    assert np.array_equal(g["mean"][b], o["mean"][b]), label + ": boundary values differ"
    out = (g["location"] == 0) | (g["location"] == 3)
    assert np.all(np.isnan(g["mean"][out]))


def test_annulus_fp64_matches_oracle(gpu_ctx, oracle):
    mesh = M.annulus()
    om = oracle.mesh(*mesh)
    eq = Pr.annulus_laplace()
    g, o = _run_both(gpu_ctx, om, mesh, eq, (64, 64, 1), 256, 1e-3, 12345)
    _compare_fp64(g, o, "annulus fp64")
    ins = g["location"] == 1
    assert np.mean(g["mean_steps"][ins] == o["mean_steps"][ins]) >= 0.97
    s = g["summary"]
    assert s["inside_count"] == ins.sum()
    assert s["total_walks"] == ins.sum() * 256


def test_annulus_fp32_tracks_fp64_and_analytic(gpu_ctx):
    V, F, B = M.annulus()
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.annulus_laplace()
    n = 1024
    a = gpu_ctx.solve_grid(64, 64, 1, eq, Context.params(n, 1e-3, 12345, precision=WOS_FP64_PLAIN))
    # exact FP32 queries: the same walks as FP64 up to rounding (common
    # random numbers), so the means agree far inside the standard errors
    gpu_ctx.set_tuning(relax=0.0)
    c = gpu_ctx.solve_grid(64, 64, 1, eq, Context.params(n, 1e-3, 12345,
                                                         precision=WOS_FP32_PERSISTENT))
    # default FP32 (relaxed queries, distance grid): where a query is relaxed
    # the walk takes a smaller sphere and becomes an independent estimate of
    # the same expectation (here nearly every grid cell holds a primitive
    # list, which answers exactly, so most walks stay the FP64 ones)
    gpu_ctx.set_tuning()
    b = gpu_ctx.solve_grid(64, 64, 1, eq, Context.params(n, 1e-3, 12345,
                                                         precision=WOS_FP32_PERSISTENT))
    assert np.array_equal(a["location"], b["location"])
    assert np.array_equal(a["location"], c["location"])
    ins = a["location"] == 1
    d = np.abs(a["mean"][ins] - c["mean"][ins])
    comb = np.sqrt(a["standard_error"][ins] ** 2 + c["standard_error"][ins] ** 2)
    print(f"fp32 exact vs fp64: max|d|={d.max():.3e} mean|d|={d.mean():.3e} "
          f"max d/se={np.max(d[comb > 0] / comb[comb > 0]):.3f}")
    assert np.all(d <= 3.0 * comb)
    comb = np.sqrt(a["standard_error"][ins] ** 2 + b["standard_error"][ins] ** 2)
    d = a["mean"][ins] - b["mean"][ins]
    assert np.all(np.abs(d[comb == 0]) <= 1e-12)  # every walk ends at once: both exact
    z = d[comb > 0] / comb[comb > 0]
    print(f"fp32 default vs fp64: max|z|={np.abs(z).max():.3f} mean z^2={np.mean(z ** 2):.3f}")
    assert np.all(np.abs(z) <= 5.0)
    assert np.mean(z ** 2) < 1.4
    # exact solution u = ln r / ln 0.25; bias budget: epsilon shell and the 128-gon
    xs = np.linspace(-1, 1, 64)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2).ravel()
    exact = np.log(r[ins]) / np.log(0.25)
    # 2904 points at 1024 walks: a distributional check on the z-scores of
    # FP32 and FP64 against u, with a bias budget of 2e-3 (epsilon shell
    # and 128-gon sagitta, both below 1e-3 here) taken off each error
    for label, res in (("fp32", b), ("fp64", a)):
        err = res["mean"][ins] - exact
        se = res["standard_error"][ins]
        live = se > 0
        zz = np.maximum(np.abs(err[live]) - 2e-3, 0.0) / se[live]
        k = int(np.argmax(np.abs(err)))
        print(f"{label} vs exact: max err={np.abs(err).max():.3e} (se {se[k]:.3e}, "
              f"u {exact[k]:.4f}) mean err={np.abs(err).mean():.3e} max z={zz.max():.2f} "
              f"mean z^2={np.mean(zz ** 2):.3f} |z|>3: {np.mean(zz > 3):.4f}")
        # SE = 0: every walk ended at once (start within epsilon: g at the
        # nearest point, off u by at most |grad u| eps = eps / (0.25 ln 4)),
        # or all n walks exited on one circle (likely only where u or 1 - u
        # is below a few / n: P = (1 - p)^n, e^-8 at p = 8 / n)
        assert np.all(np.abs(err[~live]) <= 4e-3 + 8.0 / n)
        assert zz.max() <= 5.0 and np.mean(zz ** 2) <= 1.2 and np.mean(zz > 3) <= 0.01


def test_icosphere_codes_and_fp64(gpu_ctx, oracle):
    mesh = M.icosphere(3)
    om = oracle.mesh(*mesh)
    eq = Pr.sphere_screened_constant()
    g, o = _run_both(gpu_ctx, om, mesh, eq, (16, 16, 16), 64, 1e-4, 7)
    assert np.array_equal(g["ambiguous_ray_retries"], o["ambiguous"])
    _compare_fp64(g, o, "icosphere screened fp64")


CASES_2D = [
    ("laplace2d", lambda: M.annulus(64), lambda: Pr.laplace_builtin(2)),
    ("poisson2d-uniform", lambda: M.circle(128), lambda: Pr.poisson_builtin(2, 0)),
    ("poisson2d-green", lambda: M.circle(128), lambda: Pr.poisson_builtin(2, 1)),
    ("screened2d-uniform", lambda: M.circle(128), lambda: Pr.screened_builtin(2, 5.0, 0)),
    ("screened2d-green", lambda: M.circle(128), lambda: Pr.screened_builtin(2, 5.0, 1)),
    ("screened2d-green-strong", lambda: M.circle(96), lambda: Pr.screened_builtin(2, 40.0, 1)),
    ("gaussian2d-green", lambda: M.star(512, 0.3, 5), lambda: Pr.star_poisson_gaussian(1)),
]
CASES_3D = [
    ("laplace3d", lambda: M.icosphere(2), lambda: Pr.laplace_builtin(3)),
    ("poisson3d-zero", lambda: M.icosphere(2), lambda: Pr.poisson_builtin(3, 0)),
    ("screened3d", lambda: M.icosphere(2), lambda: Pr.screened_builtin(3, 5.0)),
    ("screened3d-green", lambda: M.icosphere(2), lambda: Pr.sphere_screened_constant(10.0, 1.0, 1)),
    ("screened3d-green-strong", lambda: M.icosphere(2), lambda: Pr.sphere_screened_constant(60.0, 1.0, 1)),
    ("poisson3d-gauss-green", lambda: M.icosphere(2),
     lambda: Pr.plate_gaussian(alpha=1e-9, q=1.0, sigma=0.5, x0=(0.2, 0.0, 0.0), source_mode=1)),
]


@pytest.mark.parametrize("name,mesh_fn,eq_fn", CASES_2D + CASES_3D, ids=[c[0] for c in CASES_2D + CASES_3D])
def test_equations_fp64_vs_oracle_and_fp32(gpu_ctx, oracle, name, mesh_fn, eq_fn):
    mesh = mesh_fn()
    om = oracle.mesh(*mesh)
    eq = eq_fn()
    dim = mesh[1].shape[1]
    grid = (14, 14, 1) if dim == 2 else (8, 8, 8)
    eps = 1e-3 if dim == 2 else 1e-3
    g, o = _run_both(gpu_ctx, om, mesh, eq, grid, 48, eps, 99)
    _compare_fp64(g, o, name, exact_frac=0.9)
    f = gpu_ctx.solve_grid(*grid, eq, Context.params(48, eps, 99, precision=WOS_FP32_PERSISTENT))
    assert np.array_equal(f["location"], g["location"])
    # no rejection sampler gave up (screened radius samplers, Green modes)
    assert g["summary"]["failed_walks"] == 0 and f["summary"]["failed_walks"] == 0
    ins = g["location"] == 1
    d = np.abs(f["mean"][ins] - g["mean"][ins])
    comb = np.sqrt(f["standard_error"][ins] ** 2 + g["standard_error"][ins] ** 2) + 1e-12
    print(f"{name} fp32-fp64: max|d|={d.max():.3e} max d/se={np.max(d / comb):.3f}")
    assert np.all(d <= 4.0 * comb + 1e-6)


def test_vertex_values_torus(gpu_ctx, oracle):
    V, F, _ = M.torus(1.0, 0.35, 40, 20)
    vals = np.random.default_rng(3).standard_normal(len(V))
    eq = Pr.laplace_vertex_values(vals)
    om = oracle.mesh(V, F)
    g, o = _run_both(gpu_ctx, om, (V, F, None), eq, (10, 10, 4), 32, 1e-3, 3)
    _compare_fp64(g, o, "torus vertex values", exact_frac=0.9)


def test_point_list_edge_cases(gpu_ctx, oracle):
    V, F, B = M.annulus(32)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.annulus_laplace()
    pts = np.array([[0.5, 0.1], [5.0, 5.0], [0.0, 0.0], list(V[0]), list(V[40]), [0.6, -0.2]])
    for prec in (WOS_FP64_PLAIN, WOS_FP32_PERSISTENT):
        r = gpu_ctx.solve_points(pts, eq, Context.params(1, 1e-3, 5, precision=prec))
        assert list(r["location"]) == [1, 0, 0, 2, 2, 1]
        assert r["variance"][0] == 0.0 and r["standard_error"][0] == 0.0
        assert r["mean"][3] == 1.0 and r["mean"][4] == 0.0
        r = gpu_ctx.solve_points(pts, eq, Context.params(16, 1e-3, 5, max_steps=1, precision=prec))
        # every walk takes exactly one step; it counts as a max-steps hit
        # unless that step landed inside the epsilon shell
        assert r["mean_steps"][0] == 1.0 and r["max_steps_hits"][0] >= 10
    om = oracle.mesh(V, F, B)
    o = om.solve_points(eq, pts, 64, 1e-3, seed=5, rng_mode=1, n_threads=1)
    r = gpu_ctx.solve_points(pts, eq, Context.params(64, 1e-3, 5))
    assert np.array_equal(r["location"], o["location"])
    ins = r["location"] == 1
    assert np.allclose(r["mean"][ins], o["mean"][ins], rtol=0, atol=1e-12)


@pytest.mark.parametrize("precision", [WOS_FP32_PERSISTENT, WOS_FP64_PLAIN])
def test_bitwise_reproducible(gpu_ctx, precision):
    # both walkers hand out walks to lanes dynamically (atomic counters), so
    # two runs schedule them differently; a walk's result must not depend on it
    V, F, _ = M.icosphere(3)
    gpu_ctx.set_mesh(V, F)
    eq = Pr.sphere_screened_constant()
    p = Context.params(128, 1e-4, 7, precision=precision)
    a = gpu_ctx.solve_grid(12, 12, 12, eq, p)
    b = gpu_ctx.solve_grid(12, 12, 12, eq, p)
    for k in ("mean", "variance", "mean_steps", "location"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("shape", ["tetrahedron", "octahedron", "box12"])
def test_classification_on_shallow_trees(gpu_ctx, oracle, shape):
    """Meshes whose BVH has leaves above the start query's split depth (a
    root leaf, leaves at depth 1 and 2): location codes, start distances and
    FP64 means equal the oracle's."""
    if shape == "tetrahedron":
        V = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
        F = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], np.int32)
    elif shape == "octahedron":
        V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
        F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5],
                      [0, 3, 5]], np.int32)
    else:
        V, F, _ = M.box(1.0, 0.7, 0.4, 1, 1, 1)
    mesh = (V, F, None)
    om = oracle.mesh(*mesh)
    eq = Pr.laplace_builtin(3)
    g, o = _run_both(gpu_ctx, om, mesh, eq, (9, 9, 9), 32, 1e-3, 17)
    _compare_fp64(g, o, shape, exact_frac=0.9)
    assert np.array_equal(g["ambiguous_ray_retries"], o["ambiguous"])


@pytest.mark.parametrize("shape", ["triangle", "rectangle", "octagon", "star"])
def test_2d_classification_ties_and_shallow_trees(gpu_ctx, oracle, shape):
    """2D meshes whose BVH has leaves above the split start query's depth
    (3, 4 and 8 segments) and grid points at equal distance from two or more
    segments (corners, diagonals, the centre).  Alternating boundary ids make
    a boundary point's value depend on which segment the search picks:
    location codes, boundary values and FP64 means equal the oracle's."""
    if shape == "triangle":
        V = np.array([[-1.0, -0.5], [1.0, -0.5], [0.0, 1.0]])
    elif shape == "rectangle":
        V = np.array([[-1.0, -0.5], [1.0, -0.5], [1.0, 0.5], [-1.0, 0.5]])
    elif shape == "octagon":
        V = M.polygon_loop(8, lambda t: np.full_like(t, 1.0))
    else:
        V = M.polygon_loop(96, lambda t: 0.6 + 0.3 * np.cos(5 * t))
    n = len(V)
    F = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1).astype(np.int32)
    B = (np.arange(n) % 2).astype(np.int32)
    mesh = (V.astype(np.float64), F, B)
    om = oracle.mesh(*mesh)
    eq = Pr.laplace_builtin(2)
    g, o = _run_both(gpu_ctx, om, mesh, eq, (9, 9, 1), 64, 1e-3, 23)
    _compare_fp64(g, o, shape, exact_frac=0.9)
    assert (g["location"] == 2).any(), "the grid should hold boundary points"


def test_empty_outside_and_single_axis_inputs(gpu_ctx, oracle):
    """An empty point list is refused (the reference's grids have at least
    one point, src/main.cpp:233); a list with no inside point returns its
    location codes and NaN statistics without walking; a grid with a single
    point along an axis places it at the axis minimum
    (include/wos/grid.hpp:13-18) and equals the oracle."""
    V, F, B = M.annulus(32)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.annulus_laplace()
    for prec in (WOS_FP64_PLAIN, WOS_FP32_PERSISTENT):
        with pytest.raises(RuntimeError):
            gpu_ctx.solve_points(np.zeros((0, 2)), eq, Context.params(8, 1e-3, 5, precision=prec))
        r = gpu_ctx.solve_points(np.array([[5.0, 5.0], [0.0, 0.0], [-3.0, 0.1]]), eq,
                                 Context.params(8, 1e-3, 5, precision=prec))
        assert list(r["location"]) == [0, 0, 0]
        assert np.all(np.isnan(r["mean"])) and np.all(np.isnan(r["mean_steps"]))
    om = oracle.mesh(V, F, B)
    for grid in ((1, 9, 1), (9, 1, 1), (1, 1, 1)):
        g = gpu_ctx.solve_grid(*grid, eq, Context.params(64, 1e-3, 5))
        o = om.solve_grid(eq, *grid, 64, 1e-3, seed=5, rng_mode=1, n_threads=1)
        assert np.array_equal(g["location"], o["location"]), grid
        assert np.array_equal(g["mean"], o["mean"], equal_nan=True) or np.allclose(
            g["mean"], o["mean"], rtol=0, atol=1e-12, equal_nan=True), grid
    V3, F3, _ = M.icosphere(2)
    gpu_ctx.set_mesh(V3, F3)
    eq3 = Pr.sphere_screened_constant()
    om3 = oracle.mesh(V3, F3)
    g = gpu_ctx.solve_grid(1, 5, 3, eq3, Context.params(32, 1e-4, 9))
    o = om3.solve_grid(eq3, 1, 5, 3, 32, 1e-4, seed=9, rng_mode=1, n_threads=1)
    assert np.array_equal(g["location"], o["location"])
    assert np.allclose(g["mean"], o["mean"], rtol=0, atol=1e-12, equal_nan=True)


def test_deep_points_match_oracle(gpu_ctx, oracle):
    """Points near the centre of the icosphere are about equally far from
    every triangle, so their start queries run past the subtree budget of the
    first classification pass and are searched again on a whole wavefront
    (64 subtrees): codes and FP64 means equal the oracle's."""
    V, F, B = M.icosphere(4)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.sphere_screened_constant(10.0, 1.0)
    pts = np.random.default_rng(4).uniform(-0.25, 0.25, size=(64, 3))
    o = oracle.mesh(V, F, B).solve_points(eq, pts, 16, 1e-3, seed=5, rng_mode=1, n_threads=1)
    r = gpu_ctx.solve_points(pts, eq, Context.params(16, 1e-3, 5))
    assert np.array_equal(r["location"], o["location"])
    assert np.all(r["location"] == 1)
    assert np.allclose(r["mean"], o["mean"], rtol=0, atol=1e-12)
