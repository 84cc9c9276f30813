"""PyTorch autograd over the walk-on-spheres estimator.

`wos_points(ctx, points, eq, params, vertex_values, theta)` returns the
per-point estimates u as a tensor.  Its backward calls the adjoint entry
point of libwos_hip.so (`wos_adjoint_points`), which replays the forward
walks from their counter streams and differentiates them pathwise; nothing
of the forward walks is stored.

Differentiable inputs:
  vertex_values : [n_verts] per-vertex Dirichlet data (WOS_BND_VERTEX problems)
  theta         : [6] = (alpha, src_value, src_radius, src_center x, y, z)
Tensors may live on any device; the library exchanges host buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .engine import Context


def _descriptor(base: capi.EqDesc, theta, vertex_values):
    e = capi.EqDesc()
    C.pointer(e)[0] = base
    if theta is not None:
        t = [float(x) for x in theta]
        e.alpha, e.src_value, e.src_radius = t[0], t[1], t[2]
        for k in range(3):
            e.src_center[k] = t[3 + k]
    if vertex_values is not None:
        vv = np.ascontiguousarray(vertex_values, np.float64)
        e.vertex_values = capi.ptr(vv, C.c_double)
        e.n_vertex_values = vv.size
        e._keep = vv
    return e


class _WoSPoints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertex_values, theta, state):
        eng, pts, base, prm = state
        vv = None if vertex_values is None else vertex_values.detach().cpu().double().numpy()
        th = None if theta is None else theta.detach().cpu().double().numpy()
        eq = _descriptor(base, th, vv)
        out = eng.solve_points(pts, eq, prm)
        ctx.state = state
        ctx.vv, ctx.th = vv, th
        ctx.has_vv = vertex_values is not None
        ctx.has_th = theta is not None
        ctx.device = (vertex_values if vertex_values is not None else theta).device \
            if (vertex_values is not None or theta is not None) else torch.device("cpu")
        return torch.from_numpy(out["mean"]).to(ctx.device)

    @staticmethod
    def backward(ctx, grad_u):
        eng, pts, base, prm = ctx.state
        eq = _descriptor(base, ctx.th, ctx.vv)
        g = grad_u.detach().cpu().double().numpy().copy()
        g[~np.isfinite(g)] = 0.0
        _, gv, gp = eng.adjoint_points(pts, eq, prm, g)
        d_vv = torch.from_numpy(gv).to(ctx.device) if ctx.has_vv else None
        d_th = torch.from_numpy(gp[:6].copy()).to(ctx.device) if ctx.has_th else None
        return d_vv, d_th, None


def wos_points(ctx: Context, points, eq: capi.EqDesc, params: capi.RunParams,
               vertex_values: torch.Tensor | None = None, theta: torch.Tensor | None = None):
    """u(points) with gradients to vertex_values and theta (see module doc).
    Points outside the domain give NaN; their incoming gradient is ignored."""
    pts = np.ascontiguousarray(points, np.float64)
    return _WoSPoints.apply(vertex_values, theta, (ctx, pts, eq, params))
