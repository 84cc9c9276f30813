"""FP64 on the reference's own random streams (precision WOS_FP64_REFERENCE).

Each point's walks draw from per-point xoshiro256++ streams seeded as in
include/wos/runner.hpp:208-221 (2D: a walk and a source stream; 3D: one
shared stream), one walk after the other, so every walk is the reference's
walk, step for step.  Checked against
* per-walk values the compiled reference wrote (tests/golden/walks.npz, 15
  equation variants), through the statistics of a point seeded like the
  fixture's: point_seed(99, 4242) is the seed of list index 4242 of a point
  list solved with seed 99;
* the reference's sweep statistics (tests/golden/sweeps.npz) and the
  oracle's sweeps in its reference stream mode.
OCML and glibc transcendentals can differ in the last bit, so values agree to
rounding (1e-13 relative; 1e-11 where the reference takes its 2D screened
weights from libstdc++'s Bessel functions) while step counts agree exactly; a
walk that such a difference moved across a termination test would change the
draws of the point's later walks, which the thresholds leave room for."""
import importlib.util
import os

import numpy as np
import pytest

from wos_amd import meshgen as M
from wos_amd import problems as Pr
from wos_amd.capi import WOS_FP64_REFERENCE
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden_cases():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.eq_cases(), mg.meshes()


def _welford(values, steps):
    """Sequential Welford in walk order, as include/wos/wos.hpp:169-180."""
    n, mu, m2 = 0.0, 0.0, 0.0
    for x in values:
        n += 1.0
        d = x - mu
        mu += d / n
        m2 += d * (x - mu)
    var = m2 / (n - 1.0) if n > 1 else 0.0
    return mu, var, float(np.sum(steps)) / n


def test_walks_equal_reference_golden(gpu_ctx):
    g = np.load(os.path.join(GOLD, "walks.npz"))
    cases, meshes = _golden_cases()
    exact = 0
    for name, mname, eqf, p0, n, eps in cases:
        V, F, B = meshes[mname]
        gpu_ctx.set_mesh(V, F, B)
        pts = np.tile(np.array(p0, float), (4243, 1))
        out = gpu_ctx.solve_points(pts, eqf(), Context.params(n, eps, 99, max_steps=1000,
                                                              precision=WOS_FP64_REFERENCE))
        assert int(g[f"{name}__point_seed"][0]) != 0
        mu, var, msteps = _welford(g[f"{name}__values"], g[f"{name}__steps"])
        gm, gv, gs = out["mean"][4242], out["variance"][4242], out["mean_steps"][4242]
        exact += gm == mu and gv == var
        # the reference evaluates the 2D screened weights with libstdc++'s
        # Bessel functions, the GPU with the oracle's series (1e-12 apart);
        # elsewhere only last-bit transcendental differences remain
        tol = 1e-11 if name.startswith("screened2d") else 1e-13
        print(f"{name}: gpu mean {gm:.17g} ref {mu:.17g} steps {gs} / {msteps}")
        assert gs == msteps, name
        assert abs(gm - mu) <= tol * max(1.0, abs(mu)), name
        assert abs(gv - var) <= 1e-9 * max(1e-12, abs(var)), name
    print(f"bit-identical statistics in {exact} of {len(cases)} cases")


@pytest.mark.parametrize("case", ["annulus_laplace", "circle_poisson_green", "ico_screened"])
def test_sweeps_match_reference_golden(gpu_ctx, case):
    g = np.load(os.path.join(GOLD, "sweeps.npz"))
    cases = {"annulus_laplace": (M.annulus(32), Pr.laplace_builtin(2), (12, 12, 1), 32, 1e-3),
             "circle_poisson_green": (M.circle(48), Pr.poisson_builtin(2, 1), (10, 10, 1), 32, 1e-3),
             "ico_screened": (M.icosphere(2), Pr.screened_builtin(3, 5.0), (6, 6, 6), 16, 1e-3)}
    (V, F, B), eq, grid, n, eps = cases[case]
    gpu_ctx.set_mesh(V, F, B)
    r = gpu_ctx.solve_grid(*grid, eq, Context.params(n, eps, 31337, precision=WOS_FP64_REFERENCE))
    assert np.array_equal(r["location"], g[f"{case}__location"])
    ins = r["location"] == 1
    ref = g[f"{case}__mean"][ins]
    d = np.abs(r["mean"][ins] - ref)
    close = d <= 1e-12 * np.maximum(1.0, np.abs(ref))
    same_steps = r["mean_steps"][ins] == g[f"{case}__mean_steps"][ins]
    se = np.maximum(g[f"{case}__standard_error"][ins], 1e-300)
    print(f"{case}: inside {ins.sum()} means to 1e-12 {close.mean():.3f} "
          f"bit-identical {np.mean(d == 0):.3f} equal step counts {same_steps.mean():.3f} "
          f"max |dmean| {d.max():.3e}")
    assert close.mean() >= 0.9, case
    assert same_steps.mean() >= 0.9, case
    assert np.all(d <= 3.0 * se + 1e-12), case
    rv = np.abs(r["variance"][ins] - g[f"{case}__variance"][ins])
    assert np.mean(rv <= 1e-9 * np.maximum(1e-12, g[f"{case}__variance"][ins])) >= 0.9, case
    b = r["location"] == 2
    assert np.array_equal(r["mean"][b], g[f"{case}__mean"][b])


def test_sweep_matches_oracle_reference_mode(gpu_ctx, oracle):
    mesh = M.star(96, 0.3, 5)
    eq = Pr.star_poisson_gaussian(1)
    gpu_ctx.set_mesh(*mesh)
    r = gpu_ctx.solve_grid(24, 24, 1, eq, Context.params(64, 1e-3, 7, precision=WOS_FP64_REFERENCE))
    o = oracle.mesh(*mesh).solve_grid(eq, 24, 24, 1, 64, 1e-3, seed=7, rng_mode=0, n_threads=16)
    assert np.array_equal(r["location"], o["location"])
    ins = r["location"] == 1
    same = np.abs(r["mean"][ins] - o["mean"][ins]) <= 1e-11 * np.maximum(1.0, np.abs(o["mean"][ins]))
    se = np.maximum(o["standard_error"][ins], 1e-300)
    print(f"star gaussian green: inside {ins.sum()} exact-to-1e-11 {same.mean():.3f}")
    assert same.mean() >= 0.9
    assert np.all(np.abs(r["mean"][ins] - o["mean"][ins]) <= 3.0 * se + 1e-11)


def test_streams_continue_across_rounds(gpu_ctx):
    """Adaptive rounds split a point's walks; its streams continue, so a point
    that ends with W walks has the statistics of a fixed solve with W walks."""
    V, F, B = M.icosphere(2)
    gpu_ctx.set_mesh(V, F, B)
    eq = Pr.sphere_screened_constant(10.0, 1.0, 0)
    ad = gpu_ctx.solve_grid(6, 6, 6, eq, Context.params(16, 1e-3, 3, precision=WOS_FP64_REFERENCE,
                                                        target_se=2e-4, max_walks=256))
    assert ad["summary"]["rounds"] >= 2
    ins = ad["location"] == 1
    counts = np.unique(ad["walks"][ins])
    checked = 0
    for w in counts[:4]:
        fx = gpu_ctx.solve_grid(6, 6, 6, eq, Context.params(int(w), 1e-3, 3, precision=WOS_FP64_REFERENCE))
        sel = ins & (ad["walks"] == w)
        for k in ("mean", "variance", "mean_steps"):
            assert np.array_equal(ad[k][sel], fx[k][sel]), (w, k)
        checked += int(sel.sum())
    assert checked > 0


def test_adjoint_refuses_reference_streams(gpu_ctx):
    V, F, B = M.icosphere(1)
    gpu_ctx.set_mesh(V, F, B)
    pts = np.zeros((4, 3))
    with pytest.raises(RuntimeError):
        gpu_ctx.adjoint_points(pts, Pr.sphere_screened_constant(10.0, 1.0, 0),
                               Context.params(8, 1e-3, 1, precision=WOS_FP64_REFERENCE), np.ones(4))
