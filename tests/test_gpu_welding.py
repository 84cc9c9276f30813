"""Welding transients through the C ABI: the field source (WOS_SRC_FIELD)
and the implicit-Euler time loops of wos_amd.welding.

* One step on synthetic previous fields (level 0 bilinear, level 2 cubic with
  the clamp and a switched Gaussian, level 3 cubic with two moving terms):
  FP64 plain against the oracle's counter streams, to rounding.  The
  reference's own walks of the same three steps are golden fixtures
  (tests/golden/walks.npz, checked by the reference-stream tests).
* Level 0: the WoS centre value against the discrete implicit-Euler decay
  1 / (1 + 2 pi^2 kappa dt)^n (level0_cooling.cpp:286-288, :394-404).
* Levels 2 and 3: the field after the heating phase against the
  semi-discrete modal solution (exact in time steps, a sine series in space),
  the kind of reference apps/welding/tools/validate_level2.py builds; the
  remaining difference is the interpolation of the history on the grid.
"""
import importlib.util
import os

import numpy as np
import pytest

from wos_amd import welding as W
from wos_amd.capi import WOS_FP32_PERSISTENT, WOS_FP64_PLAIN
from wos_amd.engine import Context

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


@pytest.mark.parametrize("level", [0, 2, 3])
def test_field_step_fp64_matches_oracle(gpu_ctx, oracle, level):
    mg = _golden()
    name, mname, eqf, _, _, eps = mg._field_cases()[[0, 2, 3].index(level)]
    mesh = mg.meshes()[mname]
    eq = eqf()
    gpu_ctx.set_mesh(*mesh)
    grid = (25, 17, 1)
    g = gpu_ctx.solve_grid(*grid, eq, Context.params(96, eps, 4242, precision=WOS_FP64_PLAIN))
    o = oracle.mesh(*mesh).solve_grid(eq, *grid, 96, eps, seed=4242, rng_mode=1, n_threads=16)
    assert np.array_equal(g["location"], o["location"])
    ins = g["location"] == 1
    assert ins.sum() == 23 * 15
    dm = np.abs(g["mean"][ins] - o["mean"][ins])
    scale = np.maximum(1.0, np.abs(o["mean"][ins]))
    close = dm <= 1e-11 * scale
    se = np.maximum(o["standard_error"][ins], 1e-300)
    print(f"{name}: exact-to-1e-11 {close.mean():.4f} max|dmean| {dm.max():.3e} "
          f"max dmean/se {np.max(dm / se):.3f}")
    assert close.mean() >= 0.97
    assert np.all(dm <= 3.0 * se + 1e-11 * scale)
    assert np.mean(g["mean_steps"][ins] == o["mean_steps"][ins]) >= 0.97


@pytest.mark.parametrize("precision", [WOS_FP64_PLAIN, WOS_FP32_PERSISTENT])
def test_level0_tracks_the_discrete_decay(gpu_ctx, precision):
    steps, final = W.run_level0(gpu_ctx, W.Level0Config(), precision)
    for s in steps:
        # the bilinear history loses up to ~1 % of the mode per step
        tol = 4.0 * s["center_standard_error"] + 0.012 * s["step"] * s["discrete_amplitude"]
        print(f"step {s['step']}: centre {s['center_mean']:.6f} exact {s['discrete_amplitude']:.6f} "
              f"se {s['center_standard_error']:.2e} rmse {s['rmse']:.2e}")
        assert s["center_error"] <= tol
        assert s["rmse"] <= 0.05 * s["discrete_amplitude"]
        assert s["max_steps_hits"] == 0
    assert final.shape == (17, 17) and np.all(final[0] == 0.0) and np.all(final[:, -1] == 0.0)


def _sine_basis(lo, hi, nodes, m):
    return np.sin(np.outer(np.pi * (nodes - lo) / (hi - lo), m))


def _gauss_coeffs(lo, hi, centres, sigma, m, nq=20001):
    """(2/L) int exp(-(x - c)^2 / (2 sigma^2)) sin(m pi (x - lo) / L) dx per centre."""
    xq = np.linspace(lo, hi, nq)
    s = np.sin(np.outer(m, np.pi * (xq - lo) / (hi - lo)))
    g = np.exp(-0.5 * ((xq[None, :] - np.asarray(centres)[:, None]) / sigma) ** 2)
    return np.trapezoid(g[:, None, :] * s[None, :, :], xq, axis=2) * 2.0 / (hi - lo)


def _modal(c, nx, ny, q_hat_of_step, steps, modes):
    """Semi-discrete implicit Euler: u_k = (u_{k-1} + kappa dt q_k / k_c) /
    (1 + kappa dt lambda_mn) per Dirichlet sine mode of the plate."""
    x0, x1, y0, y1 = -0.5 * c.length_x, 0.5 * c.length_x, -0.5 * c.length_y, 0.5 * c.length_y
    m = np.arange(1, modes + 1)
    kappa = c.conductivity / (c.density * c.specific_heat)
    lam = np.pi ** 2 * ((m[:, None] / (x1 - x0)) ** 2 + (m[None, :] / (y1 - y0)) ** 2)
    u = np.zeros((modes, modes))
    for k in range(1, steps + 1):
        u = (u + kappa * c.dt * q_hat_of_step(k) / c.conductivity) / (1.0 + kappa * c.dt * lam)
    xs = np.array([W.grid_coordinate(x0, x1, i, nx) for i in range(nx)])
    ys = np.array([W.grid_coordinate(y0, y1, j, ny) for j in range(ny)])
    return _sine_basis(x0, x1, xs, m) @ u @ _sine_basis(y0, y1, ys, m).T


def _compare(field, ref, label):
    peak = ref.max()
    d = field - ref
    nrmse = np.sqrt(np.mean(d ** 2)) / peak
    ci, cj = ref.shape[0] // 2, ref.shape[1] // 2
    print(f"{label}: peak {field.max():.3f} modal {peak:.3f} nrmse {nrmse:.4f} "
          f"max|d|/peak {np.abs(d).max() / peak:.4f}")
    assert nrmse <= 0.01
    assert np.abs(d).max() <= 0.06 * peak
    return ci, cj


def test_level2_run_matches_the_modal_solution(gpu_ctx):
    """The whole level-2 run (100 steps): heating to 3 s, cooling to 10 s."""
    c = W.Level2Config()
    fields = []
    rec, field = W.run_level2(gpu_ctx, c, WOS_FP32_PERSISTENT, fields=fields)
    assert len(rec) == 100 and rec[-1]["cumulative_input_energy"] == pytest.approx(
        c.electrical_power * c.efficiency * c.heat_on_time)
    assert all(r["max_steps_hits"] == 0 for r in rec)
    modes = 160
    m = np.arange(1, modes + 1)
    x0, x1, y0, y1 = -0.06, 0.06, -0.04, 0.04
    X = _gauss_coeffs(x0, x1, [0.0], c.sigma_x, m)[0]
    Y = _gauss_coeffs(y0, y1, [0.0], c.sigma_y, m)[0]
    amp = c.electrical_power * c.efficiency / (
        c.thickness * W.gaussian_integral(x0, x1, 0.0, c.sigma_x) * W.gaussian_integral(y0, y1, 0.0, c.sigma_y))
    q = np.outer(X, Y) * amp
    for steps in (30, 100):
        ref = _modal(c, c.nx, c.ny,
                     lambda k: W.heating_fraction((k - 1) * c.dt, k * c.dt, c.heat_on_time) * q,
                     steps, modes)
        ci, cj = _compare(fields[steps - 1], ref, f"level 2 at {steps / 10:.0f} s")
        s = rec[steps - 1]
        assert abs(s["center_temperature_rise"] - ref[ci, cj]) <= 0.02 * ref[ci, cj] + 4 * s["center_standard_error"]
        # the plate holds at most the energy put in (heat leaves through the edges)
        assert 0.3 * s["cumulative_input_energy"] < s["thermal_energy_rise"] < 1.02 * s["cumulative_input_energy"]
    assert np.array_equal(fields[-1], field)


def test_level3_moving_source_matches_the_modal_solution(gpu_ctx):
    """The whole level-3 run (150 steps) against the modal solution at 2 s
    (source mid-pass), 5 s (source off) and 15 s (cooled)."""
    c = W.Level3Config()
    fields = []
    rec, probes, field = W.run_level3(gpu_ctx, c, WOS_FP32_PERSISTENT, fields=fields)
    assert len(probes) == 150 and "centreline_0" in probes[-1]
    assert all(r["max_steps_hits"] == 0 for r in rec)
    assert rec[-1]["cumulative_input_energy"] == pytest.approx(
        c.electrical_power * c.efficiency * c.heat_off_time)
    modes = 160
    m = np.arange(1, modes + 1)
    x0, x1, y0, y1 = -0.06, 0.06, -0.04, 0.04
    Y = _gauss_coeffs(y0, y1, [c.weld_y], c.sigma_y, m)[0]
    qs = []
    for k in range(1, 151):
        _, _, terms = W.moving_step_terms(c, (k - 1) * c.dt, k * c.dt)
        if not terms:
            qs.append(np.zeros((modes, modes)))
            continue
        X = _gauss_coeffs(x0, x1, [t[0] for t in terms], c.sigma_x, m)
        qs.append(sum(t[4] * np.outer(X[i], Y) for i, t in enumerate(terms)))
    for steps in (50, 150):
        _compare(fields[steps - 1], _modal(c, c.nx, c.ny, lambda k: qs[k - 1], steps, modes),
                 f"level 3 at {steps / 10:.0f} s")
    steps = 20
    field = fields[steps - 1]
    rec = rec[:steps]
    ref = _modal(c, c.nx, c.ny, lambda k: qs[k - 1], steps, modes)
    _compare(field, ref, "level 3 at 2 s")
    # the hottest node trails the source (at x = -10 mm at 2 s) where the
    # modal solution has it, within two grid cells
    s = rec[-1]
    k = int(np.argmax(ref))
    px = W.grid_coordinate(x0, x1, k // c.ny, c.nx)
    print(f"peak at x = {s['peak_x']:.4f}, modal {px:.4f}")
    assert px < -0.01 and abs(s["peak_x"] - px) <= 2.5 * 0.12 / (c.nx - 1)
    assert abs(s["peak_y"]) <= 0.0015


@pytest.mark.parametrize("case", ["level0", "level2", "level3"])
def test_transient_reference_streams_equal_reference_golden(gpu_ctx, case):
    """Small level-0/2/3 runs on the reference's streams: every step's
    field equals the one the compiled reference wrote (screened 2D: its
    libstdc++ Bessel functions and last-bit libm differences, 1e-11)."""
    from wos_amd.capi import WOS_FP64_REFERENCE
    mg = _golden()
    g = np.load(os.path.join(GOLD, "transient.npz"))
    name, c, steps = [t for t in mg.transient_cases() if t[0] == case][0]
    run = {"level0": W.run_level0, "level2": W.run_level2, "level3": W.run_level3}[case]
    fields = []
    run(gpu_ctx, c, WOS_FP64_REFERENCE, steps=steps, fields=fields)
    assert len(fields) == steps
    for k, f in enumerate(fields, 1):
        ref = g[f"{case}__step{k}"]
        d = np.abs(f - ref)
        print(f"{case} step {k}: max |d| {d.max():.3e} of {np.abs(ref).max():.4g}, "
              f"bit-identical {np.mean(d == 0):.3f}")
        assert np.allclose(f, ref, rtol=1e-11, atol=1e-13), (case, k)


def test_level2_run_on_reference_streams_equals_reference(gpu_ctx):
    """The whole level-2 run (100 steps, 1457 points; 100 walks) on the
    reference's streams: every step's centre and peak rise and the final
    field equal the run the compiled reference made step by step
    (tests/golden/transient_level2_full.npz), to last-bit rounding carried
    through 100 steps."""
    from wos_amd.capi import WOS_FP64_REFERENCE
    g = np.load(os.path.join(GOLD, "transient_level2_full.npz"))
    fields = []
    rec, final = W.run_level2(gpu_ctx, W.Level2Config(walks=100), WOS_FP64_REFERENCE,
                              fields=fields)
    centre = np.array([r["center_temperature_rise"] for r in rec])
    peak = np.array([f.max() for f in fields])
    d = np.abs(final - g["final"])
    print(f"final field: max |d| {d.max():.3e} of {g['final'].max():.4g}, "
          f"bit-identical {np.mean(d == 0):.3f}; centre max |d| {np.abs(centre - g['centre']).max():.3e}")
    assert np.allclose(final, g["final"], rtol=1e-10, atol=1e-12)
    assert np.allclose(centre, g["centre"], rtol=1e-10, atol=1e-12)
    assert np.allclose(peak, g["peak"], rtol=1e-10, atol=1e-12)
