"""ctypes bindings of the C oracle (liboracle.so) and of the reference core
driver (_ref/libwosref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libwosref.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class OrcNpq(C.Structure):
    _fields_ = [("distance", C.c_double), ("point", C.c_double * 3), ("prim", C.c_int),
                ("bid", C.c_int), ("s", C.c_double), ("t", C.c_double)]


def build(ref: bool = False):
    subprocess.check_call(["make", "-s", "-C", HERE, "all"] + (["ref"] if ref else []))


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        L = self.L = C.CDLL(path)
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_xoshiro_next.restype = C.c_uint64
        L.orc_xoshiro_seed.argtypes = [_U64, C.c_uint64]
        L.orc_xoshiro_next.argtypes = [_U64]
        L.orc_philox4x32_10.argtypes = [_U32, _U32, _U32]
        L.orc_counter_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _D]
        L.orc_point_seed.restype = C.c_uint64
        L.orc_point_seed.argtypes = [C.c_uint64, C.c_uint64]
        for name in ("orc_i0_scaled", "orc_k0_scaled", "orc_log_i0", "orc_log_k0"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.orc_bessel_j0_series.restype = C.c_double
        L.orc_bessel_j0_series.argtypes = [C.c_double]
        for name in ("orc_poisson_green_2d", "orc_poisson_green_3d", "orc_screened_weight_2d",
                     "orc_screened_weight_3d", "orc_screened_mass_2d", "orc_screened_mass_3d",
                     "orc_helmholtz_weight_2d", "orc_helmholtz_weight_3d"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double, C.c_double]
        for name in ("orc_screened_green_2d", "orc_screened_green_3d"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_npq_seg.restype = C.c_double
        L.orc_npq_seg.argtypes = [_D, _D, _D, _D, _D]
        L.orc_eberly.restype = C.c_double
        L.orc_eberly.argtypes = [_D, _D, _D, _D, _D, _D]
        L.orc_moller_trumbore.argtypes = [_D, _D, _D, _D, _D]
        L.orc_bvh_build.restype = C.c_void_p
        L.orc_bvh_build.argtypes = [C.c_int, _D, C.c_int, _I, C.c_int, _I]
        L.orc_bvh_free.argtypes = [C.c_void_p]
        L.orc_bvh_n_nodes.argtypes = [C.c_void_p]
        L.orc_bvh_export.argtypes = [C.c_void_p, _I, _I, _D, _D, _I]
        L.orc_bvh_npq.argtypes = [C.c_void_p, _D, C.POINTER(OrcNpq), _I64, _I64]
        L.orc_npq_brute.argtypes = [C.c_void_p, _D, C.POINTER(OrcNpq)]
        L.orc_mesh_bbox.argtypes = [C.c_void_p, _D]
        L.orc_winding_number.argtypes = [C.c_void_p, _D]
        L.orc_ray_cast.argtypes = [C.c_void_p, _D, _D, _I]
        L.orc_classify.argtypes = [C.c_void_p, _D, C.c_double, C.c_double, C.c_int,
                                   C.c_uint64, _I]
        L.orc_wos_run.argtypes = [C.c_void_p, C.c_void_p, _D, C.c_int, C.c_double, C.c_int,
                                  C.c_int, C.c_uint64, _D, _I, _D, _I]
        L.orc_solve_points.argtypes = [C.c_void_p, C.c_void_p, _D, C.c_int64, _I64, C.c_int,
                                       C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                       C.c_int, _D, _D, _D, _D, _U8, _I32, _I32, _I64]
        for name in ("orc_field_sample", "orc_field_source_value"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_count_reset.argtypes = [C.c_int]
        L.orc_count_get.argtypes = [_I64]
        L.orc_solve_grid.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_double, C.c_int, C.c_int, C.c_uint64,
                                     C.c_int, C.c_int, _D, _D, _D, _D, _U8, _I32, _I32, _I64,
                                     _D, _D]

    # ---- rng
    def splitmix64(self, x):
        return self.L.orc_splitmix64(x)

    def xoshiro(self, seed, n):
        s = np.zeros(4, np.uint64)
        self.L.orc_xoshiro_seed(_p(s, _U64), seed)
        return np.array([self.L.orc_xoshiro_next(_p(s, _U64)) for _ in range(n)], np.uint64)

    def philox(self, ctr, key):
        c = np.asarray(ctr, np.uint32)
        k = np.asarray(key, np.uint32)
        o = np.zeros(4, np.uint32)
        self.L.orc_philox4x32_10(_p(c, _U32), _p(k, _U32), _p(o, _U32))
        return o

    def counter_draws(self, stream_seed, walk, step, first, n):
        out = np.zeros(n)
        self.L.orc_counter_draws(stream_seed, walk, step, first, n, _p(out, _D))
        return out

    # ---- geometry
    def npq_seg(self, s0, s1, p):
        s0, s1, p = (np.asarray(x, np.float64) for x in (s0, s1, p))
        n = np.zeros(2)
        par = np.zeros(1)
        d = self.L.orc_npq_seg(_p(s0, _D), _p(s1, _D), _p(p, _D), _p(n, _D), _p(par, _D))
        return d, n, par[0]

    def eberly(self, p, v0, v1, v2):
        p, v0, v1, v2 = (np.asarray(x, np.float64) for x in (p, v0, v1, v2))
        n = np.zeros(3)
        st = np.zeros(2)
        d = self.L.orc_eberly(_p(p, _D), _p(v0, _D), _p(v1, _D), _p(v2, _D), _p(n, _D),
                              _p(st, _D))
        return d, n, st

    def moller_trumbore(self, p, ray, a, b, c):
        args = [np.asarray(x, np.float64) for x in (p, ray, a, b, c)]
        return self.L.orc_moller_trumbore(*[_p(x, _D) for x in args])

    def mesh(self, verts, prims, bids=None):
        return OracleMesh(self, verts, prims, bids)

    def count_reset(self, on=True):
        self.L.orc_count_reset(1 if on else 0)

    def count_get(self):
        o = np.zeros(3, np.int64)
        self.L.orc_count_get(_p(o, _I64))
        return dict(node_reads=int(o[0]), prim_tests=int(o[1]), queries=int(o[2]))


class OracleMesh:
    def __init__(self, orc: Oracle, verts, prims, bids=None):
        self.o = orc
        self.verts = np.ascontiguousarray(verts, np.float64)
        self.prims = np.ascontiguousarray(prims, np.int32)
        self.dim = self.prims.shape[1]
        self.bids = None if bids is None else np.ascontiguousarray(bids, np.int32)
        self.h = orc.L.orc_bvh_build(self.dim, _p(self.verts, _D), len(self.verts),
                                     _p(self.prims, _I), len(self.prims), _p(self.bids, _I))

    def __del__(self):
        try:
            self.o.L.orc_bvh_free(self.h)
        except Exception:
            pass

    def export(self):
        n = self.o.L.orc_bvh_n_nodes(self.h)
        pc = np.zeros(n, np.int32)
        idx = np.zeros(n, np.int32)
        bmin = np.zeros((n, 3))
        bmax = np.zeros((n, 3))
        leaf = np.zeros(len(self.prims), np.int32)
        self.o.L.orc_bvh_export(self.h, _p(pc, _I), _p(idx, _I), _p(bmin, _D), _p(bmax, _D),
                                _p(leaf, _I))
        return dict(prim_count=pc, idx=idx, bmin=bmin, bmax=bmax, prims=leaf)

    def bbox(self):
        g = np.zeros(6)
        self.o.L.orc_mesh_bbox(self.h, _p(g, _D))
        return g

    def npq(self, p, brute=False, counts=False):
        q = np.ascontiguousarray(p, np.float64)
        r = OrcNpq()
        nr = np.zeros(1, np.int64)
        pt = np.zeros(1, np.int64)
        if brute:
            self.o.L.orc_npq_brute(self.h, _p(q, _D), C.byref(r))
        else:
            self.o.L.orc_bvh_npq(self.h, _p(q, _D), C.byref(r), _p(nr, _I64), _p(pt, _I64))
        out = dict(distance=r.distance, point=np.array(r.point[:self.dim]), prim=r.prim,
                   bid=r.bid, s=r.s, t=r.t)
        if counts:
            out["node_reads"] = int(nr[0])
            out["prim_tests"] = int(pt[0])
        return out

    def winding(self, p):
        q = np.ascontiguousarray(p, np.float64)
        return self.o.L.orc_winding_number(self.h, _p(q, _D))

    def classify(self, p, distance, tol, max_ray_attempts, inside_seed):
        q = np.ascontiguousarray(p, np.float64)
        amb = C.c_int(0)
        code = self.o.L.orc_classify(self.h, _p(q, _D), distance, tol, max_ray_attempts,
                                     inside_seed, C.byref(amb))
        return code, amb.value

    def walks(self, eq, p0, n_walks, eps, max_steps, rng_mode, point_seed):
        q = np.ascontiguousarray(p0, np.float64)
        st = np.zeros(4)
        hits = C.c_int(0)
        vals = np.zeros(n_walks)
        steps = np.zeros(n_walks, np.int32)
        self.o.L.orc_wos_run(self.h, C.addressof(eq), _p(q, _D), n_walks, eps, max_steps,
                             rng_mode, point_seed, _p(st, _D), C.byref(hits), _p(vals, _D),
                             _p(steps, _I))
        return dict(mean=st[0], variance=st[1], standard_error=st[2], mean_steps=st[3],
                    hits=hits.value, values=vals, steps=steps)

    def solve_grid(self, eq, nx, ny, nz, n_walks, eps, max_steps=1000, max_ray_attempts=12,
                   seed=0, rng_mode=1, n_threads=8):
        if self.dim == 2:
            nz = 1
        n = nx * ny * nz
        out = _alloc(n)
        ts = np.zeros(1, np.int64)
        g = np.zeros(6)
        tol = np.zeros(1)
        self.o.L.orc_solve_grid(self.h, C.addressof(eq), nx, ny, nz, n_walks, eps, max_steps,
                                max_ray_attempts, seed, rng_mode, n_threads,
                                *_out_ptrs(out), _p(ts, _I64), _p(g, _D), _p(tol, _D))
        out["total_steps"] = int(ts[0])
        out["grid_bounds"] = g
        out["tol"] = float(tol[0])
        return out

    def solve_points(self, eq, pts, n_walks, eps, max_steps=1000, max_ray_attempts=12, seed=0,
                     rng_mode=1, n_threads=8, seed_index=None):
        pts = np.ascontiguousarray(pts, np.float64)
        n = len(pts)
        out = _alloc(n)
        ts = np.zeros(1, np.int64)
        si = None if seed_index is None else np.ascontiguousarray(seed_index, np.int64)
        self.o.L.orc_solve_points(self.h, C.addressof(eq), _p(pts, _D), n, _p(si, _I64), n_walks,
                                  eps, max_steps, max_ray_attempts, seed, rng_mode, n_threads,
                                  *_out_ptrs(out), _p(ts, _I64))
        out["total_steps"] = int(ts[0])
        return out


def _alloc(n):
    return dict(mean=np.zeros(n), variance=np.zeros(n), standard_error=np.zeros(n),
                mean_steps=np.zeros(n), location=np.zeros(n, np.uint8),
                hits=np.zeros(n, np.int32), ambiguous=np.zeros(n, np.int32))


def _out_ptrs(out):
    return (_p(out["mean"], _D), _p(out["variance"], _D), _p(out["standard_error"], _D),
            _p(out["mean_steps"], _D), _p(out["location"], _U8), _p(out["hits"], _I32),
            _p(out["ambiguous"], _I32))


class Reference:
    """The reference's own compiled core (oracle/_ref/libwosref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.L = C.CDLL(path)
        L.ref_splitmix64.restype = C.c_uint64
        L.ref_splitmix64.argtypes = [C.c_uint64]
        L.ref_prng_draws.argtypes = [C.c_uint64, C.c_int, _U64]
        L.ref_npq_seg.restype = C.c_double
        L.ref_npq_seg.argtypes = [_D, _D, _D, _D]
        L.ref_eberly.restype = C.c_double
        L.ref_eberly.argtypes = [_D, _D, _D, _D, _D]
        L.ref_moller_trumbore.argtypes = [_D, _D, _D, _D, _D]
        L.ref_mesh_create.restype = C.c_void_p
        L.ref_mesh_create.argtypes = [C.c_int, _D, C.c_int, _I, C.c_int, _I]
        L.ref_mesh_free.argtypes = [C.c_void_p]
        L.ref_bvh_n_nodes.argtypes = [C.c_void_p]
        L.ref_bvh_export.argtypes = [C.c_void_p, _I, _I, _D, _D, _I]
        L.ref_bvh_npq.restype = C.c_double
        L.ref_bvh_npq.argtypes = [C.c_void_p, _D, _D, _I, _I]
        L.ref_winding_number.argtypes = [C.c_void_p, _D]
        L.ref_walks.argtypes = [C.c_void_p, C.c_void_p, _D, C.c_int, C.c_double, C.c_int,
                                C.c_uint64, _D, _I]
        L.ref_solve_points.argtypes = [C.c_void_p, C.c_void_p, _D, _I64, C.c_int64, C.c_int,
                                       C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                       _D, _D, _D, _D, _U8, _I32, _I32]

    def prng(self, seed, n):
        o = np.zeros(n, np.uint64)
        self.L.ref_prng_draws(seed, n, _p(o, _U64))
        return o

    def mesh(self, verts, prims, bids=None):
        return RefMesh(self, verts, prims, bids)


class RefMesh:
    def __init__(self, ref: Reference, verts, prims, bids=None):
        self.r = ref
        self.verts = np.ascontiguousarray(verts, np.float64)
        self.prims = np.ascontiguousarray(prims, np.int32)
        self.dim = self.prims.shape[1]
        self.bids = None if bids is None else np.ascontiguousarray(bids, np.int32)
        self.h = ref.L.ref_mesh_create(self.dim, _p(self.verts, _D), len(self.verts),
                                       _p(self.prims, _I), len(self.prims), _p(self.bids, _I))

    def __del__(self):
        try:
            self.r.L.ref_mesh_free(self.h)
        except Exception:
            pass

    def export(self):
        n = self.r.L.ref_bvh_n_nodes(self.h)
        pc = np.zeros(n, np.int32)
        idx = np.zeros(n, np.int32)
        bmin = np.zeros((n, 3))
        bmax = np.zeros((n, 3))
        leaf = np.zeros(len(self.prims), np.int32)
        self.r.L.ref_bvh_export(self.h, _p(pc, _I), _p(idx, _I), _p(bmin, _D), _p(bmax, _D),
                                _p(leaf, _I))
        return dict(prim_count=pc, idx=idx, bmin=bmin, bmax=bmax, prims=leaf)

    def npq(self, p):
        q = np.ascontiguousarray(p, np.float64)
        n = np.zeros(3)
        prim = C.c_int(0)
        bid = C.c_int(0)
        d = self.r.L.ref_bvh_npq(self.h, _p(q, _D), _p(n, _D), C.byref(prim), C.byref(bid))
        return dict(distance=d, point=n[:self.dim], prim=prim.value, bid=bid.value)

    def winding(self, p):
        q = np.ascontiguousarray(p, np.float64)
        return self.r.L.ref_winding_number(self.h, _p(q, _D))

    def walks(self, eq, p0, n_walks, eps, max_steps, point_seed):
        q = np.ascontiguousarray(p0, np.float64)
        vals = np.zeros(n_walks)
        steps = np.zeros(n_walks, np.int32)
        self.r.L.ref_walks(self.h, C.addressof(eq), _p(q, _D), n_walks, eps, max_steps,
                           point_seed, _p(vals, _D), _p(steps, _I))
        return vals, steps

    def solve_points(self, eq, pts, n_walks, eps, max_steps=1000, max_ray_attempts=12, seed=0,
                     n_threads=8, seed_index=None):
        pts = np.ascontiguousarray(pts, np.float64)
        n = len(pts)
        out = _alloc(n)
        si = None if seed_index is None else np.ascontiguousarray(seed_index, np.int64)
        self.r.L.ref_solve_points(self.h, C.addressof(eq), _p(pts, _D), _p(si, _I64), n,
                                  n_walks, eps, max_steps, max_ray_attempts, seed, n_threads,
                                  *_out_ptrs(out))
        return out


def grid_points(bounds, nx, ny, nz, dim):
    """Grid coordinates exactly as include/wos/grid.hpp:12-18."""
    def coord(lo, hi, i, n):
        return lo + (hi - lo) * float(i) / float(n - 1) if n > 1 else lo
    if dim == 2:
        nz = 1
    pts = np.zeros((nx * ny * nz, dim))
    for i in range(nx):
        x = coord(bounds[0], bounds[1], i, nx)
        for j in range(ny):
            y = coord(bounds[2], bounds[3], j, ny)
            for k in range(nz):
                lin = (i * ny + j) * nz + k
                pts[lin, 0] = x
                pts[lin, 1] = y
                if dim == 3:
                    pts[lin, 2] = coord(bounds[4], bounds[5], k, nz)
    return pts
