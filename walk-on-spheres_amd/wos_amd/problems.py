"""Equation descriptors: the reference's built-in equations and the
benchmark problems.

Built-ins restate the data of src/equations/laplace.cpp:12-43,
poisson.cpp:15-118 and screened_poisson.cpp:40-141 as wos_eq_desc values.
"""
from __future__ import annotations

from .capi import (WOS_BND_ID_CONST, WOS_BND_ONE, WOS_BND_SIN3X_COS2Y, WOS_BND_VERTEX,
                   WOS_BND_X2Y2_ID0, WOS_BND_Y5, WOS_BND_ZERO, WOS_EQ_HELMHOLTZ,
                   WOS_EQ_LAPLACE, WOS_EQ_POISSON, WOS_EQ_SCREENED,
                   WOS_SOURCE_GREEN, WOS_SOURCE_GREEN_MIS, WOS_SOURCE_UNIFORM, WOS_SRC_CONSTANT, WOS_SRC_DISK_SQ,
                   WOS_SRC_DISK_SQRT, WOS_SRC_GAUSSIAN, WOS_SRC_NONE, WOS_SRC_ZERO, make_eq)


# ---- reference built-ins (the `wos` registry entries)
def laplace_builtin(dim: int):
    """Laplace2D: g = x^2 - y^2 on boundary id 0 (laplace.cpp:16-18);
    Laplace3D: g = y5^0 + 4 y5^-3 (laplace.cpp:30-36)."""
    return make_eq(WOS_EQ_LAPLACE, boundary_kind=WOS_BND_X2Y2_ID0 if dim == 2 else WOS_BND_Y5)


def poisson_builtin(dim: int, source_mode: int = WOS_SOURCE_UNIFORM):
    """Poisson2D: f = 0.1 in the disk r <= 0.15 at the origin, g = 0
    (poisson.cpp:22-48); Poisson3D: f = 0 (still sampled), g = Y5
    (poisson.cpp:79-88)."""
    if dim == 2:
        return make_eq(WOS_EQ_POISSON, source_kind=WOS_SRC_DISK_SQ, boundary_kind=WOS_BND_ZERO,
                       source_mode=source_mode, src_value=0.1, src_center=(0.0, 0.0, 0.0),
                       src_radius=0.15)
    return make_eq(WOS_EQ_POISSON, source_kind=WOS_SRC_ZERO, boundary_kind=WOS_BND_Y5,
                   source_mode=source_mode)


def screened_builtin(dim: int, alpha: float = 5.0, source_mode: int = WOS_SOURCE_UNIFORM):
    """ScreenedPoisson2D: f = 5 in the disk |x-(0.45,0)| <= 0.5, g = x^2-y^2 on
    id 0 (screened_poisson.cpp:48-65); ScreenedPoisson3D: no source, g = Y5
    (screened_poisson.cpp:94-112)."""
    if dim == 2:
        return make_eq(WOS_EQ_SCREENED, source_kind=WOS_SRC_DISK_SQRT,
                       boundary_kind=WOS_BND_X2Y2_ID0, source_mode=source_mode, alpha=alpha,
                       src_value=5.0, src_center=(0.45, 0.0, 0.0), src_radius=0.5)
    return make_eq(WOS_EQ_SCREENED, source_kind=WOS_SRC_NONE, boundary_kind=WOS_BND_Y5,
                   source_mode=source_mode, alpha=alpha)


def helmholtz_builtin(dim: int, k: float = 1.0):
    """Helmholtz2D/3D: Delta u + k^2 u = 0 with k = 1, no source, g = 1
    (helmholtz.cpp:15-78); k travels in the descriptor's alpha."""
    del dim
    return make_eq(WOS_EQ_HELMHOLTZ, source_kind=WOS_SRC_NONE, boundary_kind=WOS_BND_ONE, alpha=k)


BUILTINS = {"laplace": laplace_builtin, "poisson": poisson_builtin,
            "screened_poisson": screened_builtin, "helmholtz": helmholtz_builtin}


# ---- benchmark problems
def annulus_laplace():
    """2D Laplace, g = 1 on the inner circle (boundary id 0), 0 on the outer;
    exact u = ln r / ln 0.25."""
    return make_eq(WOS_EQ_LAPLACE, boundary_kind=WOS_BND_ID_CONST, bnd_values=(1.0, 0.0))


def sphere_screened_constant(alpha: float = 10.0, f: float = 1.0,
                             source_mode: int = WOS_SOURCE_UNIFORM):
    """3D screened Poisson, constant source, g = 0; on the unit ball
    u = (f/alpha^2) (1 - sinh(alpha r) / (r sinh alpha))."""
    return make_eq(WOS_EQ_SCREENED, source_kind=WOS_SRC_CONSTANT, boundary_kind=WOS_BND_ZERO,
                   source_mode=source_mode, alpha=alpha, src_value=f)


def star_poisson_gaussian(source_mode: int = WOS_SOURCE_GREEN):
    """2D Poisson, f = exp(-|x|^2 / 0.1), g = sin(3x) cos(2y)."""
    return make_eq(WOS_EQ_POISSON, source_kind=WOS_SRC_GAUSSIAN,
                   boundary_kind=WOS_BND_SIN3X_COS2Y, source_mode=source_mode, src_value=1.0,
                   src_center=(0.0, 0.0, 0.0), src_radius=0.1 ** 0.5)


def laplace_vertex_values(values):
    """Laplace with per-vertex Dirichlet data blended on the nearest primitive."""
    return make_eq(WOS_EQ_LAPLACE, boundary_kind=WOS_BND_VERTEX, vertex_values=values)


def plate_gaussian(alpha: float = 0.05, q: float = 1.0, sigma: float = 3.0,
                   x0=(40.0, 30.0, 10.0), source_mode: int = WOS_SOURCE_UNIFORM):
    """3D screened Poisson on the plate, f = Q exp(-|x-x0|^2/sigma^2), g = 0."""
    return make_eq(WOS_EQ_SCREENED, source_kind=WOS_SRC_GAUSSIAN, boundary_kind=WOS_BND_ZERO,
                   source_mode=source_mode, alpha=alpha, src_value=q, src_center=tuple(x0),
                   src_radius=sigma)


# ---- welding level 1: stationary Gaussian heat in a thin plate
WELDING_L1 = dict(length_x=0.12, length_y=0.08, thickness=0.006, conductivity=35.0,
                  electrical_power=800.0, efficiency=0.75, sigma=0.006, nx=49, ny=33,
                  walks=6000, max_steps=1000, epsilon=8e-6)


def welding_gaussian_integral(lo: float, hi: float, c: float, sigma: float) -> float:
    """sigma sqrt(pi/2) (erf((hi - c)/(sqrt2 sigma)) - erf((lo - c)/(sqrt2 sigma)))
    (apps/welding/include/welding/gaussian_heat_source.hpp:116-126)."""
    from math import erf, sqrt
    return sigma * 1.25331413731550025120788264240552263 * (
        erf((hi - c) / (sqrt(2.0) * sigma)) - erf((lo - c) / (sqrt(2.0) * sigma)))


def welding_gaussian(source_mode: int = WOS_SOURCE_GREEN, beta: float = 0.5, cfg=None,
                     truncate: bool = True, alpha: float = 0.0):
    """The level-1 welding heat problem (apps/welding/src/level1_stationary_heat.cpp:
    42-62, :624-666): 2D Poisson on the plate [0, Lx] x [0, Ly], g = 0, and
    f = q'''/k with the finite-plate Gaussian q''' = eta P exp(-|x - c|^2 / (2 sigma^2))
    / (t Ix Iy) centred on the plate (gaussian_heat_source.hpp:38-41).  In
    GREEN_MIS mode the proposal is that Gaussian truncated to the plate (or,
    truncate=False, on the whole plane).  alpha > 0 makes it the screened
    equation with the same source."""
    c = dict(WELDING_L1, **(cfg or {}))
    cx, cy, s = 0.5 * c["length_x"], 0.5 * c["length_y"], c["sigma"]
    ix = welding_gaussian_integral(0.0, c["length_x"], cx, s)
    iy = welding_gaussian_integral(0.0, c["length_y"], cy, s)
    peak = c["electrical_power"] * c["efficiency"] / (c["thickness"] * ix * iy)
    box = (0.0, c["length_x"], 0.0, c["length_y"]) if truncate else (0.0, 0.0, 0.0, 0.0)
    return make_eq(WOS_EQ_SCREENED if alpha > 0.0 else WOS_EQ_POISSON,
                   source_kind=WOS_SRC_GAUSSIAN, boundary_kind=WOS_BND_ZERO,
                   source_mode=source_mode, alpha=alpha, src_value=peak / c["conductivity"],
                   src_center=(cx, cy, 0.0), src_radius=s * 2.0 ** 0.5,
                   mis_green_probability=beta if source_mode == WOS_SOURCE_GREEN_MIS else 0.0,
                   proposal_box=box)
