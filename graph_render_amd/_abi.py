"""ctypes binding of libgraph_render_fa2.so (include/graph_render_fa2.h).

The library is required: every product entry point goes through it, and loading fails
loudly when it has not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgraph_render_fa2.so")

GR_OK = 0
GR_ERR_CAPACITY = 1
GR_ERR_NON_FINITE = 2
GR_ERR_PARAM = 3
GR_ERR_SNAPSHOT = 4
GR_ERR_INVALID_HANDLE = 10
GR_ERR_UNKNOWN_LAYOUT = 11
GR_ERR_PARAM_OUT_OF_RANGE = 20
GR_ERR_PARAMS_MALFORMED = 21
GR_ERR_DEVICE = 30
GR_ERR_UNSUPPORTED = 31

GR_PREC_FP64, GR_PREC_FP32, GR_PREC_FP32_KAHAN = 0, 1, 2
GR_REP_DEFAULT, GR_REP_TREE, GR_REP_DENSE, GR_REP_PM = 0, 1, 2, 3
GR_EXT_EXPORT_TREE = 1
GR_EXT_DIRECT_LAUNCH = 2
GR_EXT_BLOCK_SUMS = 4


class GrFa2Params(C.Structure):
    _fields_ = [("max_iter", C.c_uint32), ("jitter_tolerance", C.c_double),
                ("scaling_ratio", C.c_double), ("gravity", C.c_double),
                ("seed", C.c_uint32), ("dim", C.c_uint32)]


class GrForceSimParams(C.Structure):
    """gr_forcesim_params: Fa2ForceSimParams (layout/force/forcesim.rs:104-153)."""
    _fields_ = [("iterations", C.c_uint32), ("scale", C.c_double),
                ("scaling_ratio", C.c_double), ("gravity", C.c_double),
                ("jitter_tolerance", C.c_double), ("seed", C.c_uint64)]


class GrD3Params(C.Structure):
    """gr_d3_params: ForceParams (layout/force/params.rs:26-77)."""
    _fields_ = [(f, C.c_double) for f in (
        "charge_strength", "theta", "distance_min", "distance_max", "link_distance",
        "link_strength_scale", "collide_radius", "center_strength", "velocity_decay",
        "alpha_decay", "alpha_min", "initial_alpha")]


class GrFa2Schedule(C.Structure):
    _fields_ = [("walk_block", C.c_uint32), ("walk_xcd", C.c_uint32),
                ("sort_items", C.c_uint32), ("finish_split", C.c_uint32)]


class GrFa2Ext(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("flags", C.c_uint32), ("theta", C.c_double),
                ("precision", C.c_int32), ("repulsion", C.c_int32),
                ("edge_weights", C.c_void_p), ("start", C.c_void_p),
                ("pm_grid", C.c_uint32), ("reserved", C.c_uint32)]


# every function include/graph_render_fa2.h declares, with its ctypes signature
SIGNATURES = {
    "gr_fa2_params_default": (None, [C.POINTER(GrFa2Params)]),
    "gr_fa2_ext_default": (None, [C.POINTER(GrFa2Ext)]),
    "gr_fa2_params_decode": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(GrFa2Params)]),
    "gr_layout_run": (C.c_int, [C.c_char_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                                C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "gr_fa2_session_create": (C.c_void_p, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.POINTER(GrFa2Params), C.c_void_p]),
    "gr_fa2_batch_create": (C.c_void_p, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.POINTER(GrFa2Params), C.c_void_p]),
    "gr_fa2_session_release": (None, [C.c_void_p]),
    "gr_fa2_session_step": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "gr_fa2_session_set_params": (C.c_int, [C.c_void_p, C.POINTER(GrFa2Params), C.c_double]),
    "gr_fa2_session_positions": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gr_fa2_session_set_positions": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gr_fa2_session_forces": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "gr_fa2_session_scalars": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gr_fa2_session_graph_scalars": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "gr_fa2_session_graph": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "gr_fa2_session_tree": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "gr_fa2_session_record": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_uint32)]),
    "gr_fa2_session_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    "gr_fa2_session_walk_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gr_fa2_schedule_select": (C.c_int, [C.c_uint32, C.c_size_t, C.c_uint32, C.c_int32,
                                         C.c_uint32, C.POINTER(GrFa2Schedule)]),
    "gr_fa2_session_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(GrFa2Schedule)]),
    "gr_fa2_session_set_schedule": (C.c_int, [C.c_void_p, C.POINTER(GrFa2Schedule)]),
    "gr_fa2_session_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(GrFa2Schedule)]),
    "gr_fa2_session_time": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32,
                                      C.POINTER(C.c_double)]),
    "gr_pm_side_for": (C.c_uint32, [C.c_uint32]),
    "gr_pm_charge": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "gr_pm_charge_noalloc": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint32, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_void_p,
                                       C.c_void_p]),
    "gr_d3_charge_bh": (C.c_int, [C.c_uint32, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gr_d3_charge_bh_time": (C.c_int, [C.c_uint32, C.c_void_p, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "gr_quadtree_node_ids": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint32,
                                       C.POINTER(C.c_uint32), C.c_void_p]),
    "gr_d3_collide": (C.c_int, [C.c_uint32, C.c_void_p, C.c_double, C.c_uint32, C.c_uint32,
                                C.c_void_p]),
    "gr_d3_params_default": (None, [C.c_void_p]),
    "gr_pm_layout_run": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                   C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_double)]),
    "gr_d3_session_with_particle_mesh": (C.c_int, [C.c_void_p]),
    "gr_pm_d3_charge": (C.c_int, [C.c_uint32, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_void_p, C.c_void_p]),
    "gr_pm_d3_collide": (C.c_int, [C.c_uint32, C.c_void_p, C.c_double, C.c_uint32, C.c_uint32,
                                   C.c_void_p, C.c_void_p]),
    "gr_d3_session_create": (C.c_void_p, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_size_t, C.c_void_p]),
    "gr_d3_session_step": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_int32)]),
    "gr_d3_session_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "gr_d3_session_time": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "gr_d3_session_alpha": (C.c_double, [C.c_void_p]),
    "gr_d3_session_reheat": (C.c_int, [C.c_void_p, C.c_double]),
    "gr_d3_session_set_alpha_target": (C.c_int, [C.c_void_p, C.c_double]),
    "gr_d3_session_release": (None, [C.c_void_p]),
    "gr_d3_layout_run": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                   C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_double)]),
    "gr_forcesim_params_default": (None, [C.c_void_p]),
    "gr_forcesim_run": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "gr_forcesim_batch": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "gr_forcesim_batch_time": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_uint32,
                                         C.POINTER(C.c_double)]),
    "gr_snapshot_encode": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_size_t, C.POINTER(C.c_size_t)]),
    "gr_last_error": (C.c_int, [C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]),
    "gr_version": (C.c_char_p, []),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """The loaded HIP library; raises LibraryMissing if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run graph_render_amd/build.py (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    name = C.c_char_p()
    rule = C.c_char_p()
    code = lib().gr_last_error(C.byref(name), C.byref(rule))
    return code, (name.value or b"").decode(), (rule.value or b"").decode()
