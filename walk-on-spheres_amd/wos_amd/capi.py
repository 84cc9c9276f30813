"""ctypes mirror of include/wos_hip.h and the loader of libwos_hip.so.

The structures here are byte-for-byte the C ABI declared in
include/wos_hip.h.  `load()` opens the in-tree HIP library and raises if it
is missing: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "csrc", "libwos_hip.so")

# enum wos_status
WOS_OK, WOS_ERR_INPUT, WOS_ERR_MESH, WOS_ERR_DEVICE, WOS_ERR_IO, WOS_ERR_STATE = range(6)
# enum wos_equation_kind
WOS_EQ_LAPLACE, WOS_EQ_POISSON, WOS_EQ_SCREENED, WOS_EQ_HELMHOLTZ = range(4)
# enum wos_source_mode (values of the reference SourceMode)
WOS_SOURCE_UNIFORM, WOS_SOURCE_GREEN, WOS_SOURCE_GREEN_MIS = range(3)
# enum wos_source_kind
(WOS_SRC_NONE, WOS_SRC_DISK_SQ, WOS_SRC_DISK_SQRT, WOS_SRC_CONSTANT, WOS_SRC_GAUSSIAN,
 WOS_SRC_ZERO, WOS_SRC_FIELD) = range(7)
# enum wos_boundary_kind
(WOS_BND_X2Y2_ID0, WOS_BND_Y5, WOS_BND_ZERO, WOS_BND_ID_CONST, WOS_BND_VERTEX,
 WOS_BND_SIN3X_COS2Y, WOS_BND_ONE) = range(7)
# enum wos_precision
WOS_FP64_PLAIN, WOS_FP32_PERSISTENT, WOS_FP64_REFERENCE = range(3)
WOS_MAX_BND_VALUES = 16

LOC_OUTSIDE, LOC_INSIDE, LOC_BOUNDARY, LOC_INDETERMINATE = range(4)
# enum wos_grad_index
WOS_GRAD_ALPHA, WOS_GRAD_SRC_VALUE, WOS_GRAD_SRC_RADIUS, WOS_GRAD_SRC_CENTER = 0, 1, 2, 3
WOS_GRAD_BND_VALUES = 6
WOS_N_GRAD = 6 + WOS_MAX_BND_VALUES


WOS_INTERP_BILINEAR, WOS_INTERP_CUBIC = 0, 1
WOS_MAX_HEAT_TERMS = 4


class HeatTerm(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("sx", C.c_double),
                ("sy", C.c_double), ("a", C.c_double), ("b", C.c_double)]


class FieldSource(C.Structure):
    _fields_ = [
        ("values", C.POINTER(C.c_double)),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("bounds", C.c_double * 4),
        ("interpolation", C.c_int32),
        ("clamp_nonnegative", C.c_int32),
        ("history_scale", C.c_double),
        ("heat_scale", C.c_double),
        ("heat_divisor", C.c_double),
        ("n_heat_terms", C.c_int32),
        ("reserved0", C.c_int32),
        ("heat", HeatTerm * WOS_MAX_HEAT_TERMS),
    ]


class EqDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("source_mode", C.c_int32),
        ("source_kind", C.c_int32),
        ("boundary_kind", C.c_int32),
        ("alpha", C.c_double),
        ("src_value", C.c_double),
        ("src_center", C.c_double * 3),
        ("src_radius", C.c_double),
        ("bnd_values", C.c_double * WOS_MAX_BND_VALUES),
        ("n_bnd_values", C.c_int32),
        ("reserved0", C.c_int32),
        ("vertex_values", C.POINTER(C.c_double)),
        ("n_vertex_values", C.c_int64),
        ("mis_green_probability", C.c_double),
        ("proposal_box", C.c_double * 4),
        ("field_source", C.POINTER(FieldSource)),
    ]


class RunParams(C.Structure):
    _fields_ = [
        ("n_walks", C.c_int32),
        ("max_steps", C.c_int32),
        ("max_ray_attempts", C.c_int32),
        ("precision", C.c_int32),
        ("epsilon", C.c_double),
        ("seed", C.c_uint64),
        ("target_se", C.c_double),
        ("max_walks", C.c_int32),
        ("tune", C.c_int32),
        ("count_units", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class PointStats(C.Structure):
    _fields_ = [
        ("mean", C.POINTER(C.c_double)),
        ("variance", C.POINTER(C.c_double)),
        ("standard_error", C.POINTER(C.c_double)),
        ("mean_steps", C.POINTER(C.c_double)),
        ("location", C.POINTER(C.c_uint8)),
        ("max_steps_hits", C.POINTER(C.c_int32)),
        ("walks", C.POINTER(C.c_int32)),
        ("ambiguous_ray_retries", C.POINTER(C.c_int32)),
    ]


class RunSummary(C.Structure):
    _fields_ = [
        ("inside_count", C.c_int64),
        ("max_steps_hits", C.c_int64),
        ("ambiguous_ray_retries", C.c_int64),
        ("indeterminate_points", C.c_int64),
        ("sum_variance", C.c_double),
        ("sum_standard_error", C.c_double),
        ("sum_mean_steps", C.c_double),
        ("wos_seconds", C.c_double),
        ("walk_kernel_seconds", C.c_double),
        ("total_walks", C.c_int64),
        ("total_steps", C.c_int64),
        ("grid_bounds", C.c_double * 6),
        ("boundary_tolerance", C.c_double),
        ("launch_blocks", C.c_int32),
        ("launch_threads", C.c_int32),
        ("lds_cached_nodes", C.c_int32),
        ("rounds", C.c_int32),
        ("q_queries", C.c_int64),
        ("q_grid_bound", C.c_int64),
        ("q_list_scans", C.c_int64),
        ("q_list_entries", C.c_int64),
        ("q_list_prims", C.c_int64),
        ("q_nodes", C.c_int64),
        ("q_leaf_prims", C.c_int64),
        ("failed_walks", C.c_int64),
        ("grid_build_seconds", C.c_double),
        ("grid_bytes", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("block_threads", C.c_int32),
        ("lds_nodes", C.c_int32),
        ("chunk", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("min_ready", C.c_int32),
        ("relax", C.c_float),
        ("grid_cells", C.c_int32),
        ("grid_list_max", C.c_int32),
        ("xcd_parts", C.c_int32),
        ("bvh_width", C.c_int32),
        ("stack_lds", C.c_int32),
    ]


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def make_eq(kind, source_kind=WOS_SRC_NONE, boundary_kind=WOS_BND_ZERO,
            source_mode=WOS_SOURCE_UNIFORM, alpha=0.0, src_value=0.0,
            src_center=(0.0, 0.0, 0.0), src_radius=0.0, bnd_values=(),
            vertex_values=None, mis_green_probability=0.0, proposal_box=(0.0, 0.0, 0.0, 0.0)):
    """Build a wos_eq_desc.  `vertex_values` must stay alive while in use."""
    e = EqDesc()
    e.kind = kind
    e.source_mode = source_mode
    e.source_kind = source_kind
    e.boundary_kind = boundary_kind
    e.alpha = alpha
    e.src_value = src_value
    for k in range(3):
        e.src_center[k] = src_center[k] if k < len(src_center) else 0.0
    e.src_radius = src_radius
    if len(bnd_values) > WOS_MAX_BND_VALUES:
        raise ValueError("at most %d boundary-id values" % WOS_MAX_BND_VALUES)
    for i, v in enumerate(bnd_values):
        e.bnd_values[i] = v
    e.n_bnd_values = len(bnd_values)
    e.mis_green_probability = mis_green_probability
    for k in range(4):
        e.proposal_box[k] = proposal_box[k]
    if vertex_values is not None:
        vv = np.ascontiguousarray(vertex_values, dtype=np.float64)
        e.vertex_values = ptr(vv, C.c_double)
        e.n_vertex_values = vv.size
        e._keep = vv  # keep the buffer alive with the descriptor
    return e


_LIB = None


def load(path: str | None = None):
    """Open libwos_hip.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libwos_hip.so not built at {p}; run __graft_entry__.build()")
    lib = C.CDLL(p)
    lib.wos_last_error.restype = C.c_char_p
    lib.wos_version.restype = C.c_char_p
    lib.wos_ctx_create.restype = C.c_void_p
    lib.wos_ctx_create.argtypes = [C.c_int]
    lib.wos_ctx_destroy.argtypes = [C.c_void_p]
    lib.wos_ctx_set_mesh.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int64,
                                     C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32)]
    lib.wos_solve_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(EqDesc),
                                   C.POINTER(RunParams), C.POINTER(PointStats),
                                   C.POINTER(RunSummary)]
    lib.wos_solve_points.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                                     C.POINTER(EqDesc), C.POINTER(RunParams),
                                     C.POINTER(PointStats), C.POINTER(RunSummary)]
    lib.wos_solve_points_indexed.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                             C.POINTER(C.c_int64), C.c_int64,
                                             C.POINTER(EqDesc), C.POINTER(RunParams),
                                             C.POINTER(PointStats), C.POINTER(RunSummary)]
    for name in ("wos_adjoint_points",):
        getattr(lib, name).argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                                       C.POINTER(EqDesc), C.POINTER(RunParams),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(PointStats),
                                       C.POINTER(RunSummary)]
    lib.wos_adjoint_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(EqDesc),
                                     C.POINTER(RunParams), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(PointStats), C.POINTER(RunSummary)]
    lib.wos_ctx_set_domains.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                        C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64),
                                        C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int64),
                                        C.POINTER(C.POINTER(C.c_int32))]
    lib.wos_solve_grid_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(EqDesc), C.POINTER(RunParams),
                                         C.POINTER(PointStats), C.POINTER(RunSummary),
                                         C.POINTER(RunSummary)]
    lib.wos_bvh_build_host.restype = C.c_int64
    lib.wos_bvh_build_host.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int64,
                                       C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    lib.wos_closest_points.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int64,
                                       C.c_int, C.c_double, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int32)]
    lib.wos_closest_points_noalloc.argtypes = lib.wos_closest_points.argtypes
    lib.wos_ctx_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(EqDesc),
                                     C.POINTER(RunParams), C.POINTER(Tuning)]
    lib.wos_ctx_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
    lib.wos_ctx_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
    if path is None:
        _LIB = lib
    return lib


def check(lib, rc):
    if rc != WOS_OK:
        raise RuntimeError("wos error %d: %s" % (rc, lib.wos_last_error().decode()))
