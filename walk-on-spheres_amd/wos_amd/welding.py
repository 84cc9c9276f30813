"""Implicit-Euler thin-plate heat transients on the GPU engine: the time
loops of the welding app's levels 0, 2 and 3 with every step one screened
grid solve whose source is the previous field (WOS_SRC_FIELD).

Reference counterparts:
* level 0, cooling of sin(pi x) sin(pi y) on the unit square, bilinear
  history (apps/welding/src/level0_cooling.cpp:37-47, :181-196, :262-417);
* level 2, a stationary Gaussian source switched off after heat_on_time,
  cubic history clamped at 0 (level2_transient_heat.cpp:47-76, :227-291,
  :398-640);
* level 3, a Gaussian moving along y = 0 at constant speed, time-averaged
  over each step by 2-point Gauss-Legendre (level3_moving_heat.cpp:44-76,
  :167-224, :335-600; moving_gaussian_heat_source.hpp:17-120).

Step k solves  Lap u_k - alpha^2 u_k = -(alpha^2 max(0, u_{k-1}) + q_k / k_c)
with alpha = 1 / sqrt(kappa dt), g = 0, Green-mode source sampling, on the
plate's grid; the walks of grid point `index` at step k use the per-point
seed splitmix64(seed ^ splitmix64(k) ^ splitmix64(index)) of the apps
(e.g. level2_transient_heat.cpp:578-581), which is the engine's per-point
seed for the global seed seed ^ splitmix64(k).  With precision
WOS_FP64_REFERENCE every step's walks are therefore the reference's.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import meshgen
from .capi import (WOS_BND_ZERO, WOS_EQ_SCREENED, WOS_FP64_PLAIN, WOS_INTERP_BILINEAR,
                   WOS_INTERP_CUBIC, WOS_MAX_HEAT_TERMS, WOS_SOURCE_GREEN, WOS_SRC_FIELD,
                   FieldSource, make_eq, ptr)
from .engine import Context

_M64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """include/wos/hash.hpp splitmix64 finaliser."""
    z = (x + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def step_seed(seed: int, step: int) -> int:
    """Global seed whose per-point seeds are those of time step `step`."""
    return (seed ^ splitmix64(step)) & _M64


def gaussian_integral(lo: float, hi: float, c: float, sigma: float) -> float:
    """gaussian_heat_source.hpp:116-126 / moving_gaussian_heat_source.hpp:103-112."""
    sqrt_two = 1.41421356237309504880168872420969808
    sqrt_pi_over_two = 1.25331413731550025120788264240552263
    return sigma * sqrt_pi_over_two * (math.erf((hi - c) / (sqrt_two * sigma)) -
                                       math.erf((lo - c) / (sqrt_two * sigma)))


def heating_fraction(time_begin: float, time_end: float, heat_on_time: float) -> float:
    """level2_transient_heat.cpp:227-235."""
    begin = max(0.0, time_begin)
    end = min(time_end, heat_on_time)
    if not end > begin:
        return 0.0
    return (end - begin) / (time_end - time_begin)


def grid_coordinate(lo: float, hi: float, i: int, n: int) -> float:
    """include/wos/grid.hpp:12-17."""
    return lo + (hi - lo) * float(i) / float(n - 1) if n > 1 else lo


def field_source(previous, grid, interpolation, clamp, history_scale, heat_scale=1.0,
                 heat_divisor=1.0, terms=()):
    """A wos_field_source: grid = (nx, ny, xmin, xmax, ymin, ymax); terms =
    [(cx, cy, sx, sy, a, b), ...], each a exp(-(dx^2 + dy^2)/2) / b.  The
    returned structure keeps `previous` alive."""
    fs = FieldSource()
    nx, ny, x0, x1, y0, y1 = grid
    if previous is not None:
        v = np.ascontiguousarray(previous, np.float64).reshape(-1)
        if v.size != nx * ny:
            raise ValueError("previous field must have nx * ny values")
        fs.values = ptr(v, C.c_double)
        fs._keep = v
    fs.nx, fs.ny = nx, ny
    for k, b in enumerate((x0, x1, y0, y1)):
        fs.bounds[k] = b
    fs.interpolation = interpolation
    fs.clamp_nonnegative = 1 if clamp else 0
    fs.history_scale = history_scale
    fs.heat_scale = heat_scale
    fs.heat_divisor = heat_divisor
    if len(terms) > WOS_MAX_HEAT_TERMS:
        raise ValueError("at most %d heat terms" % WOS_MAX_HEAT_TERMS)
    fs.n_heat_terms = len(terms)
    for k, t in enumerate(terms):
        h = fs.heat[k]
        h.cx, h.cy, h.sx, h.sy, h.a, h.b = t
    return fs


def field_equation(alpha: float, fs: FieldSource):
    """Screened Poisson with the field source, g = 0, Green source sampling."""
    eq = make_eq(WOS_EQ_SCREENED, source_kind=WOS_SRC_FIELD, boundary_kind=WOS_BND_ZERO,
                 source_mode=WOS_SOURCE_GREEN, alpha=alpha)
    eq.field_source = C.pointer(fs)
    eq._keep_fs = fs
    return eq


def moving_step_terms(c, time_begin, time_end):
    """MovingGaussianStepSource2D (moving_gaussian_heat_source.hpp:19-72):
    (active_fraction, centre_x, heat terms) of one step; each of the two
    Gauss-Legendre components is normalised on the finite plate."""
    xmin, xmax = -0.5 * c.length_x, 0.5 * c.length_x
    ymin, ymax = -0.5 * c.length_y, 0.5 * c.length_y
    absorbed = c.electrical_power * c.efficiency
    active_begin = max(0.0, time_begin)
    active_end = min(time_end, c.heat_off_time)
    if not active_end > active_begin:
        return 0.0, c.source_start_x + c.weld_speed * c.heat_off_time, []
    fraction = (active_end - active_begin) / (time_end - time_begin)
    midpoint = 0.5 * (active_begin + active_end)
    half_width = 0.5 * (active_end - active_begin)
    inv_sqrt3 = 0.577350269189625764509148780501957456
    terms = []
    for t in (midpoint - inv_sqrt3 * half_width, midpoint + inv_sqrt3 * half_width):
        cx = c.source_start_x + c.weld_speed * t
        if cx < xmin or cx > xmax:
            raise ValueError("moving Gaussian centre leaves the finite plate")
        ix = gaussian_integral(xmin, xmax, cx, c.sigma_x)
        iy = gaussian_integral(ymin, ymax, c.weld_y, c.sigma_y)
        coef = fraction * absorbed / (2.0 * c.thickness * ix * iy)
        terms.append((cx, c.weld_y, c.sigma_x, c.sigma_y, coef, 1.0))
    return fraction, c.source_start_x + c.weld_speed * midpoint, terms


@dataclass
class Level0Config:
    """level0_cooling.cpp:37-47."""
    grid_size: int = 17
    steps: int = 5
    walks: int = 2000
    max_steps: int = 1000
    dt: float = 0.01
    diffusivity: float = 1.0
    epsilon: float = 1e-3
    seed: int = 12345


@dataclass
class Level2Config:
    """level2_transient_heat.cpp:47-76."""
    nx: int = 49
    ny: int = 33
    walks: int = 500
    max_steps: int = 1000
    history_interpolation: int = WOS_INTERP_CUBIC
    dt: float = 0.1
    total_time: float = 10.0
    heat_on_time: float = 3.0
    epsilon: float = 8e-6
    seed: int = 20260819
    length_x: float = 0.12
    length_y: float = 0.08
    thickness: float = 0.006
    conductivity: float = 35.0
    density: float = 7800.0
    specific_heat: float = 600.0
    ambient_temperature: float = 293.15
    electrical_power: float = 800.0
    efficiency: float = 0.75
    sigma_x: float = 0.006
    sigma_y: float = 0.006


@dataclass
class Level3Config:
    """level3_moving_heat.cpp:44-76 (fixed physics, configurable walks/seed)."""
    nx: int = 97
    ny: int = 65
    max_steps: int = 1000
    dt: float = 0.1
    total_time: float = 15.0
    heat_off_time: float = 5.0
    epsilon: float = 8e-6
    length_x: float = 0.12
    length_y: float = 0.08
    thickness: float = 0.006
    conductivity: float = 35.0
    density: float = 7800.0
    specific_heat: float = 600.0
    ambient_temperature: float = 293.15
    electrical_power: float = 800.0
    efficiency: float = 0.75
    sigma_x: float = 0.006
    sigma_y: float = 0.006
    source_start_x: float = -0.05
    source_end_x: float = 0.05
    weld_y: float = 0.0
    weld_speed: float = 0.02
    walks: int = 500
    seed: int = 20260819
    probes: tuple = (("centreline_m20", -0.020, 0.000, 1.5), ("centreline_m10", -0.010, 0.000, 2.0),
                     ("centreline_0", 0.000, 0.000, 2.5), ("centreline_p10", 0.010, 0.000, 3.0),
                     ("centreline_p20", 0.020, 0.000, 3.5), ("transverse_p5", 0.000, 0.005, 2.5),
                     ("transverse_p10", 0.000, 0.010, 2.5))


def field_summary(field2d, se2d, grid, rho_cp_t):
    """summarize_field (level2_transient_heat.cpp:304-349): peak, SE
    statistics over interior nodes, trapezoidal thermal energy."""
    nx, ny, x0, x1, y0, y1 = grid
    dx, dy = (x1 - x0) / (nx - 1), (y1 - y0) / (ny - 1)
    wx = np.ones(nx)
    wx[[0, -1]] = 0.5
    wy = np.ones(ny)
    wy[[0, -1]] = 0.5
    integral = float(np.sum(wx[:, None] * wy[None, :] * field2d))
    flat = field2d.reshape(-1)
    k = int(np.argmax(flat))
    peak = max(0.0, float(flat[k]))
    pi_, pj = (k // ny, k % ny) if flat[k] > 0.0 else (0, 0)
    inner = field2d[1:-1, 1:-1]
    inner_se = se2d[1:-1, 1:-1]
    return dict(peak_rise=peak, peak_i=pi_, peak_j=pj,
                peak_x=grid_coordinate(x0, x1, pi_, nx), peak_y=grid_coordinate(y0, y1, pj, ny),
                mean_standard_error=float(inner_se.mean()) if inner.size else 0.0,
                max_standard_error=float(inner_se.max()) if inner.size else 0.0,
                minimum_interior_rise=float(inner.min()) if inner.size else 0.0,
                negative_interior_points=int(np.sum(inner < 0.0)),
                thermal_energy_rise=rho_cp_t * integral * dx * dy)


def _step(ctx, nx, ny, eq, walks, eps, seed, step, max_steps, precision):
    prm = Context.params(walks, eps, step_seed(seed, step), max_steps=max_steps,
                         precision=precision)
    t0 = time.perf_counter()
    r = ctx.solve_grid(nx, ny, 1, eq, prm)
    dt = time.perf_counter() - t0
    # boundary nodes carry g = 0; the apps leave them at 0 (never solved)
    mean = np.where(r["location"] == 1, r["mean"], 0.0).reshape(nx, ny)
    se = np.where(r["location"] == 1, r["standard_error"], 0.0).reshape(nx, ny)
    return mean, se, int(r["max_steps_hits"].sum()), dt, r


def level0_equation(c: Level0Config, prev):
    """Step equation of level 0 (level0_cooling.cpp:181-196, :361-362)."""
    alpha = 1.0 / math.sqrt(c.diffusivity * c.dt)
    n = c.grid_size
    fs = field_source(prev, (n, n, 0.0, 1.0, 0.0, 1.0), WOS_INTERP_BILINEAR, False, alpha * alpha, 0.0)
    return field_equation(alpha, fs)


def level0_initial(c: Level0Config):
    """sin(pi x) sin(pi y) on the interior nodes, 0 on the edges (:272-284)."""
    n = c.grid_size
    xs = np.array([grid_coordinate(0.0, 1.0, i, n) for i in range(n)])
    prev = np.zeros((n, n))
    prev[1:-1, 1:-1] = np.sin(math.pi * xs[1:-1, None]) * np.sin(math.pi * xs[None, 1:-1])
    return prev


def plate_grid(c):
    """The centred plate and its grid (nx, ny, xmin, xmax, ymin, ymax)."""
    return (c.nx, c.ny, -0.5 * c.length_x, 0.5 * c.length_x, -0.5 * c.length_y, 0.5 * c.length_y)


def level2_equation(c: Level2Config, prev, step: int):
    """Step equation of level 2 (level2_transient_heat.cpp:237-291, :529-545):
    returns (equation, power fraction of the step)."""
    grid = plate_grid(c)
    _, _, x0, x1, y0, y1 = grid
    kappa = c.conductivity / (c.density * c.specific_heat)
    alpha = 1.0 / math.sqrt(kappa * c.dt)
    ix = gaussian_integral(x0, x1, 0.0, c.sigma_x)
    iy = gaussian_integral(y0, y1, 0.0, c.sigma_y)
    term = (0.0, 0.0, c.sigma_x, c.sigma_y, c.electrical_power * c.efficiency,
            c.thickness * ix * iy)
    frac = heating_fraction((step - 1) * c.dt, step * c.dt, c.heat_on_time)
    fs = field_source(prev, grid, c.history_interpolation, True, alpha * alpha, frac,
                      c.conductivity, [term])
    return field_equation(alpha, fs), frac


def level3_equation(c: "Level3Config", prev, step: int):
    """Step equation of level 3 (level3_moving_heat.cpp:174-224, :496-515):
    returns (equation, active fraction of the step)."""
    grid = plate_grid(c)
    kappa = c.conductivity / (c.density * c.specific_heat)
    alpha = 1.0 / math.sqrt(kappa * c.dt)
    frac, _, terms = moving_step_terms(c, (step - 1) * c.dt, step * c.dt)
    fs = field_source(prev, grid, WOS_INTERP_CUBIC, True, alpha * alpha, 1.0, c.conductivity, terms)
    return field_equation(alpha, fs), frac


def run_level0(ctx: Context, cfg: Level0Config = None, precision=WOS_FP64_PLAIN, steps=None,
               fields=None):
    """Cooling of the sin(pi x) sin(pi y) mode on the unit square with the
    discrete decay 1 / (1 + 2 pi^2 kappa dt) per step (level0_cooling.cpp:
    286-288, :394-404).  Returns per-step records and the final field; a
    list passed as `fields` receives every step's field."""
    c = cfg or Level0Config()
    n = c.grid_size
    ctx.set_mesh(*meshgen.rectangle(0.0, 1.0, 0.0, 1.0))
    prev = level0_initial(c)
    xs = np.array([grid_coordinate(0.0, 1.0, i, n) for i in range(n)])
    decay = 1.0 / (1.0 + 2.0 * math.pi * math.pi * c.diffusivity * c.dt)
    exact_mode = np.sin(math.pi * xs[:, None]) * np.sin(math.pi * xs[None, :])
    ci = n // 2
    out = []
    for step in range(1, (steps or c.steps) + 1):
        mean, se, hits, secs, _ = _step(ctx, n, n, level0_equation(c, prev), c.walks, c.epsilon,
                                        c.seed, step, c.max_steps, precision)
        amp = decay ** step
        err = (mean - amp * exact_mode)[1:-1, 1:-1]
        out.append(dict(step=step, time=step * c.dt, center_mean=float(mean[ci, ci]),
                        center_standard_error=float(se[ci, ci]), discrete_amplitude=amp,
                        continuous_amplitude=math.exp(-2.0 * math.pi ** 2 * c.diffusivity * step * c.dt),
                        center_error=abs(float(mean[ci, ci]) - amp),
                        rmse=float(np.sqrt(np.mean(err ** 2))), max_abs_error=float(np.abs(err).max()),
                        max_steps_hits=hits, step_seconds=secs))
        if fields is not None:
            fields.append(mean)
        prev = mean
    return out, prev


def run_level2(ctx: Context, cfg: Level2Config = None, precision=WOS_FP64_PLAIN, steps=None,
               fields=None):
    """Stationary Gaussian heating for heat_on_time, then cooling
    (level2_transient_heat.cpp:398-640).  Returns per-step summaries (the
    columns of the app's summary CSV) and the final field."""
    c = cfg or Level2Config()
    grid = plate_grid(c)
    ctx.set_mesh(*meshgen.rectangle(*grid[2:]))
    n_steps = steps or int(round(c.total_time / c.dt))
    absorbed = c.electrical_power * c.efficiency
    prev = np.zeros((c.nx, c.ny))
    ci, cj = c.nx // 2, c.ny // 2
    energy_in = 0.0
    out = []
    for step in range(1, n_steps + 1):
        te = step * c.dt
        eq, frac = level2_equation(c, prev, step)
        energy_in += frac * absorbed * c.dt
        mean, se, hits, secs, _ = _step(ctx, c.nx, c.ny, eq, c.walks, c.epsilon, c.seed, step,
                                        c.max_steps, precision)
        s = field_summary(mean, se, grid, c.density * c.specific_heat * c.thickness)
        s.update(step=step, time=te, source_fraction=frac, absorbed_power=frac * absorbed,
                 cumulative_input_energy=energy_in, center_temperature_rise=float(mean[ci, cj]),
                 center_standard_error=float(se[ci, cj]), max_steps_hits=hits, step_seconds=secs)
        out.append(s)
        if fields is not None:
            fields.append(mean)
        prev = mean
    return out, prev


def run_level3(ctx: Context, cfg: Level3Config = None, precision=WOS_FP64_PLAIN, steps=None,
               fields=None):
    """Moving Gaussian source, then cooling (level3_moving_heat.cpp:335-600).
    Returns per-step summaries, per-step probe values and the final field."""
    c = cfg or Level3Config()
    grid = plate_grid(c)
    _, _, x0, x1, y0, y1 = grid
    ctx.set_mesh(*meshgen.rectangle(x0, x1, y0, y1))
    n_steps = steps or int(round(c.total_time / c.dt))
    probe_idx = []
    for name, px, py, _ in c.probes:
        i = int(round((px - x0) / (x1 - x0) * (c.nx - 1)))
        j = int(round((py - y0) / (y1 - y0) * (c.ny - 1)))
        if (abs(grid_coordinate(x0, x1, i, c.nx) - px) > 1e-12 or
                abs(grid_coordinate(y0, y1, j, c.ny) - py) > 1e-12):
            raise ValueError("a fixed Level-3 probe is not on the query grid")
        probe_idx.append((name, i, j))
    prev = np.zeros((c.nx, c.ny))
    ci, cj = c.nx // 2, c.ny // 2
    energy_in = 0.0
    out, probes = [], []
    for step in range(1, n_steps + 1):
        te = step * c.dt
        eq, frac = level3_equation(c, prev, step)
        energy_in += frac * c.electrical_power * c.efficiency * c.dt
        mean, se, hits, secs, _ = _step(ctx, c.nx, c.ny, eq, c.walks, c.epsilon, c.seed, step,
                                        c.max_steps, precision)
        s = field_summary(mean, se, grid, c.density * c.specific_heat * c.thickness)
        s.update(step=step, time=te, phase=1 if te <= c.heat_off_time else 2,
                 source_x=c.source_start_x + c.weld_speed * min(max(te, 0.0), c.heat_off_time),
                 active_fraction=frac, absorbed_power=frac * c.electrical_power * c.efficiency,
                 cumulative_input_energy=energy_in, center_temperature_rise=float(mean[ci, cj]),
                 center_standard_error=float(se[ci, cj]), max_steps_hits=hits, step_seconds=secs)
        out.append(s)
        probes.append({name: (float(mean[i, j]), float(se[i, j])) for name, i, j in probe_idx})
        if fields is not None:
            fields.append(mean)
        prev = mean
    return out, probes, prev


__all__ = ["Level0Config", "Level2Config", "Level3Config", "run_level0", "run_level2",
           "run_level3", "level0_equation", "level2_equation", "level3_equation",
           "level0_initial", "plate_grid", "field_source", "field_equation", "moving_step_terms",
           "heating_fraction", "gaussian_integral", "splitmix64", "step_seed"]
