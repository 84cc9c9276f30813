"""Python host mirror of the C ABI: a context per GPU, a mesh, grid and
point-list solves.  Everything runs in libwos_hip.so; this module only
marshals numpy arrays.

Reference counterparts: run<N,Eq> (include/wos/runner.hpp:22-308) for
`solve_grid`, wos::wos_run as the welding apps call it
(apps/welding/src/level1_stationary_heat.cpp:257) for `solve_points`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import (EqDesc, PointStats, RunParams, RunSummary, WOS_FP32_PERSISTENT,
                   WOS_FP64_PLAIN, ptr)


def screened_epsilon(epsilon: float, alpha: float) -> float:
    """Stopping distance the reference uses for screened_poisson:
    min(epsilon, 0.05 / alpha) (src/equations/screened_poisson.cpp:20-38)."""
    return min(epsilon, 0.05 / alpha)


class Context:
    """One GPU's engine state (wos_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = capi.load()
        h = self.lib.wos_ctx_create(device)
        if not h:
            raise RuntimeError("wos_ctx_create failed: " + self.lib.wos_last_error().decode())
        self.h = C.c_void_p(h)
        self.dim = None
        self.n_verts = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.wos_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, verts, prims, boundary_ids=None):
        v = np.ascontiguousarray(verts, np.float64)
        p = np.ascontiguousarray(prims, np.int32)
        b = None if boundary_ids is None else np.ascontiguousarray(boundary_ids, np.int32)
        dim = p.shape[1]
        if v.shape[1] != dim:
            v = np.ascontiguousarray(v[:, :dim])
        rc = self.lib.wos_ctx_set_mesh(self.h, dim, ptr(v, C.c_double), len(v),
                                       ptr(p, C.c_int32), len(p), ptr(b, C.c_int32))
        capi.check(self.lib, rc)
        self.dim = dim
        self.n_verts = len(v)
        self.n_domains = 1

    def set_domains(self, meshes):
        """Several domains [(verts, prims, boundary_ids or None), ...] of one
        dimension, solved together by solve_grid_batch."""
        vs = [np.ascontiguousarray(m[0], np.float64) for m in meshes]
        ps = [np.ascontiguousarray(m[1], np.int32) for m in meshes]
        dim = ps[0].shape[1]
        vs = [np.ascontiguousarray(v[:, :dim]) for v in vs]
        bs = [None if len(m) < 3 or m[2] is None else np.ascontiguousarray(m[2], np.int32)
              for m in meshes]
        n = len(meshes)
        vp = (C.POINTER(C.c_double) * n)(*[ptr(v, C.c_double) for v in vs])
        pp = (C.POINTER(C.c_int32) * n)(*[ptr(p, C.c_int32) for p in ps])
        bp = (C.POINTER(C.c_int32) * n)(*[ptr(b, C.c_int32) if b is not None else None for b in bs])
        nv = (C.c_int64 * n)(*[len(v) for v in vs])
        npr = (C.c_int64 * n)(*[len(p) for p in ps])
        rc = self.lib.wos_ctx_set_domains(self.h, dim, n, vp, nv, pp, npr, bp)
        capi.check(self.lib, rc)
        self.dim = dim
        self.n_verts = int(sum(len(v) for v in vs))
        self.n_domains = n

    def solve_grid_batch(self, nx, ny, nz, eq: EqDesc, prm: RunParams):
        """Returns (list of per-domain result dicts, total summary)."""
        if self.dim == 2:
            nz = 1
        m = nx * ny * nz
        out, st, summ = self._outputs(m * self.n_domains)
        per = (RunSummary * self.n_domains)()
        rc = self.lib.wos_solve_grid_batch(self.h, nx, ny, nz, C.byref(eq), C.byref(prm),
                                           C.byref(st), per, C.byref(summ))
        capi.check(self.lib, rc)
        res = []
        for d in range(self.n_domains):
            r = {k: v[d * m:(d + 1) * m] for k, v in out.items()}
            r["summary"] = self._summary(per[d])
            res.append(r)
        return res, self._summary(summ)

    def set_tuning(self, block_threads=128, lds_nodes=-1, chunk=64, blocks_per_cu=0,
                   min_ready=40, relax=0.2, grid_cells=-1, grid_list_max=127,
                   xcd_parts=8, bvh_width=2, stack_lds=0):
        t = capi.Tuning(block_threads, lds_nodes, chunk, blocks_per_cu, min_ready, relax,
                        grid_cells, grid_list_max, xcd_parts, bvh_width, stack_lds)
        capi.check(self.lib, self.lib.wos_ctx_set_tuning(self.h, C.byref(t)))

    def autotune(self, nx, ny, nz, eq: EqDesc, prm: RunParams):
        """Search the FP32 walker's launch shape on this grid problem (results
        are unchanged by it); returns and installs the chosen tuning."""
        if self.dim == 2:
            nz = 1
        t = capi.Tuning()
        capi.check(self.lib, self.lib.wos_ctx_autotune(self.h, nx, ny, nz, C.byref(eq),
                                                       C.byref(prm), C.byref(t)))
        return {k: getattr(t, k) for k, _ in capi.Tuning._fields_}

    def get_tuning(self):
        t = capi.Tuning()
        capi.check(self.lib, self.lib.wos_ctx_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in capi.Tuning._fields_}

    @staticmethod
    def params(n_walks, epsilon, seed, max_steps=1000, max_ray_attempts=12,
               precision=WOS_FP64_PLAIN, target_se=0.0, max_walks=0, tune=0, count_units=0,
               use_graph=0):
        p = RunParams()
        p.n_walks = n_walks
        p.max_steps = max_steps
        p.max_ray_attempts = max_ray_attempts
        p.precision = precision
        p.epsilon = epsilon
        p.seed = seed
        p.target_se = target_se
        p.max_walks = max_walks
        p.tune = tune
        p.count_units = count_units
        p.use_graph = use_graph
        return p

    def closest_points(self, points, precision=WOS_FP64_PLAIN, relax=0.0, domain=0,
                       noalloc=False):
        """Closest boundary point per point: dict(dist, lower, nearest, prim)
        (bvh_npq of the reference, include/wos/bvh.hpp:38-39).  noalloc:
        the variant that stages in the library's fixed device array
        (wos_closest_points_noalloc)."""
        dim = self.dim
        pts = np.ascontiguousarray(points, np.float64)
        n = len(pts)
        pts = np.ascontiguousarray(pts.reshape(n, -1)[:, :dim]) if n else np.zeros((0, dim))
        dist, lower = np.zeros(n), np.zeros(n)
        nearest = np.zeros((n, dim))
        prim = np.zeros(n, np.int32)
        fn = self.lib.wos_closest_points_noalloc if noalloc else self.lib.wos_closest_points
        rc = fn(self.h, domain, ptr(pts, C.c_double), n, precision, relax,
                                         ptr(dist, C.c_double), ptr(lower, C.c_double),
                                         ptr(nearest, C.c_double), ptr(prim, C.c_int32))
        capi.check(self.lib, rc)
        return dict(dist=dist, lower=lower, nearest=nearest, prim=prim)

    def adjoint_points(self, points, eq: EqDesc, prm: RunParams, dl_du):
        """Replay the walks of solve_points and return (forward stats,
        grad_vertex [n_verts], grad_params [WOS_N_GRAD])."""
        pts = np.ascontiguousarray(points, np.float64)
        return self._adjoint(self.lib.wos_adjoint_points, len(pts), eq, prm, dl_du,
                             ptr(pts, C.c_double), len(pts))

    def adjoint_grid(self, nx, ny, nz, eq: EqDesc, prm: RunParams, dl_du):
        if self.dim == 2:
            nz = 1
        return self._adjoint(self.lib.wos_adjoint_grid, nx * ny * nz, eq, prm, dl_du, nx, ny, nz)

    def _adjoint(self, fn, n, eq, prm, dl_du, *args):
        dl = np.ascontiguousarray(dl_du, np.float64)
        if dl.shape != (n,):
            raise ValueError(f"dl_du must have shape ({n},)")
        gv = np.zeros(self.n_verts)
        gp = np.zeros(capi.WOS_N_GRAD)
        out, st, summ = self._outputs(n)
        rc = fn(self.h, *args, C.byref(eq), C.byref(prm), ptr(dl, C.c_double),
                ptr(gv, C.c_double), ptr(gp, C.c_double), C.byref(st), C.byref(summ))
        capi.check(self.lib, rc)
        out["summary"] = self._summary(summ)
        return out, gv, gp

    @staticmethod
    def _summary(summ):
        s = {k: getattr(summ, k) for k, _ in RunSummary._fields_}
        s["grid_bounds"] = list(summ.grid_bounds)
        return s

    @staticmethod
    def _outputs(n):
        out = dict(mean=np.zeros(n), variance=np.zeros(n), standard_error=np.zeros(n),
                   mean_steps=np.zeros(n), location=np.zeros(n, np.uint8),
                   max_steps_hits=np.zeros(n, np.int32), walks=np.zeros(n, np.int32),
                   ambiguous_ray_retries=np.zeros(n, np.int32))
        st = PointStats()
        st.mean = ptr(out["mean"], C.c_double)
        st.variance = ptr(out["variance"], C.c_double)
        st.standard_error = ptr(out["standard_error"], C.c_double)
        st.mean_steps = ptr(out["mean_steps"], C.c_double)
        st.location = ptr(out["location"], C.c_uint8)
        st.max_steps_hits = ptr(out["max_steps_hits"], C.c_int32)
        st.walks = ptr(out["walks"], C.c_int32)
        st.ambiguous_ray_retries = ptr(out["ambiguous_ray_retries"], C.c_int32)
        return out, st, RunSummary()

    def _solve(self, fn, n, eq: EqDesc, prm: RunParams, *args):
        out = dict(mean=np.zeros(n), variance=np.zeros(n), standard_error=np.zeros(n),
                   mean_steps=np.zeros(n), location=np.zeros(n, np.uint8),
                   max_steps_hits=np.zeros(n, np.int32), walks=np.zeros(n, np.int32),
                   ambiguous_ray_retries=np.zeros(n, np.int32))
        st = PointStats()
        st.mean = ptr(out["mean"], C.c_double)
        st.variance = ptr(out["variance"], C.c_double)
        st.standard_error = ptr(out["standard_error"], C.c_double)
        st.mean_steps = ptr(out["mean_steps"], C.c_double)
        st.location = ptr(out["location"], C.c_uint8)
        st.max_steps_hits = ptr(out["max_steps_hits"], C.c_int32)
        st.walks = ptr(out["walks"], C.c_int32)
        st.ambiguous_ray_retries = ptr(out["ambiguous_ray_retries"], C.c_int32)
        summ = RunSummary()
        rc = fn(self.h, *args, C.byref(eq), C.byref(prm), C.byref(st), C.byref(summ))
        capi.check(self.lib, rc)
        out["summary"] = {k: getattr(summ, k) for k, _ in RunSummary._fields_}
        out["summary"]["grid_bounds"] = list(summ.grid_bounds)
        return out

    def solve_grid(self, nx, ny, nz, eq: EqDesc, prm: RunParams):
        if self.dim == 2:
            nz = 1
        return self._solve(self.lib.wos_solve_grid, nx * ny * nz, eq, prm, nx, ny, nz)

    def solve_points(self, points, eq: EqDesc, prm: RunParams, seed_index=None):
        """Point i is seeded by its list index, or by seed_index[i] (the
        global grid index whose seed it takes, wos_solve_points_indexed)."""
        pts = np.ascontiguousarray(points, np.float64)
        if seed_index is None:
            return self._solve(self.lib.wos_solve_points, len(pts), eq, prm,
                               ptr(pts, C.c_double), len(pts))
        si = np.ascontiguousarray(seed_index, np.int64)
        if si.shape != (len(pts),):
            raise ValueError("seed_index must have one entry per point")
        return self._solve(self.lib.wos_solve_points_indexed, len(pts), eq, prm,
                           ptr(pts, C.c_double), ptr(si, C.c_int64), len(pts))


__all__ = ["Context", "screened_epsilon", "WOS_FP64_PLAIN", "WOS_FP32_PERSISTENT"]
