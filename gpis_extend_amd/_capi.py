"""ctypes binding of the C ABI declared in include/gpis_hip.h (libgpis_hip.so).

The product path: there is no CPU fallback.  If the HIP library is missing or
cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libgpis_hip.so")

GPIS_OK, GPIS_INVALID_ARGUMENT, GPIS_DEVICE_ERROR, GPIS_UNSUPPORTED, GPIS_OUT_OF_MEMORY = range(5)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "DEVICE_ERROR", 3: "UNSUPPORTED", 4: "OUT_OF_MEMORY"}

COV_SE, COV_PROC_NS = 0, 1
MEAN_SPHERE, MEAN_BOX, MEAN_PLANE, MEAN_GRID, MEAN_PROCEDURAL = 0, 1, 2, 3, 4
# SdfFunctions::Function (SdfFunctions.hpp:13-22)
SDF_FUNCTIONS = ("knob", "knob_inner", "knob_outer", "two_spheres", "plane", "rounded_u_groove", "two_slabs",
                 "three_slabs")
CTX_GLOBAL, CTX_RENEWAL, CTX_RENEWAL_PLUS, CTX_NONE = 0, 1, 2, 3
SCHEME_UNI, SCHEME_NEE, SCHEME_MIS = 0, 1, 2
LS_CONST, LS_SINCOS, LS_NOISE = 0, 1, 2
VAR_CONST, VAR_NOISE = 0, 1
NOISE_TYPES = ("bottom_top", "left_right", "front_back", "bottom_top_left_right", "sandstone", "rust")
EXIT, HIT, FAIL, INVALID = 0, 1, 2, 3
FAIL_NONE, FAIL_OUTWARD, FAIL_ZERO_GRAD, FAIL_NONFINITE, FAIL_BOUNCE_LIMIT = 0, 1, 2, 3, 4
PHASE_MIRROR, PHASE_LAMBERT = 0, 1

#: every symbol include/gpis_hip.h declares
EXPORTS = (
    "gpis_create", "gpis_destroy", "gpis_last_error", "gpis_set_domain", "gpis_impulse_table",
    "gpis_eval_points", "gpis_trace_flights", "gpis_trace_paths", "gpis_get_counters",
    "gpis_reset_counters", "gpis_set_launch", "gpis_set_refill", "gpis_set_coop_cells", "gpis_trace_flights_1d",
    "gpis_set_mean_grid", "gpis_trace_paths_1d", "gpis_set_lists", "gpis_set_precision",
    # include/gpis_surrogate.h
    "gpis_net_create", "gpis_net_destroy", "gpis_net_last_error", "gpis_net_set_params", "gpis_net_get_params",
    "gpis_net_set_calibration", "gpis_net_features_1d", "gpis_net_classes_1d", "gpis_net_gradient",
    "gpis_net_train_step", "gpis_net_forward", "gpis_surrogate_flights_1d", "gpis_surrogate_flights_1d_noalloc",
    "gpis_net_set_averaging", "gpis_net_use_average",
)

NET_FEATURES, NET_CLASSES, NET_PARAMS = 8, 32, 54816


class FieldDesc(C.Structure):
    _fields_ = [
        ("cov_kind", C.c_int32), ("sigma", C.c_float), ("length_scale", C.c_float),
        ("kernel_scale", C.c_float), ("aniso", C.c_float * 3),
        ("ns_multires", C.c_int32), ("ns_ls_field", C.c_int32), ("ns_ls_params", C.c_float * 4),
        ("ns_ls_max", C.c_float), ("ns_var", C.c_float),
        ("mean_kind", C.c_int32), ("mean_radius", C.c_float), ("mean_center", C.c_double * 3),
        ("mean_half", C.c_double * 3), ("plane_normal", C.c_double * 3), ("plane_offset", C.c_double),
        ("global_seed", C.c_uint32), ("impulse_density", C.c_float), ("single_realization", C.c_int32),
        ("ctxt", C.c_int32), ("one_d", C.c_int32), ("scheme", C.c_int32), ("correlation_xy", C.c_int32),
        ("step_size", C.c_float), ("min_step", C.c_uint32), ("max_bounce", C.c_int32),
        ("isotropic_3d", C.c_int32),
        ("sdf_function", C.c_int32), ("sdf_inverse_transform", C.c_float * 12),
        ("sdf_min", C.c_float), ("sdf_scale", C.c_float), ("sdf_offset", C.c_float),
        ("ns_var_field", C.c_int32), ("ns_var_noise", C.c_int32), ("ns_ls_noise", C.c_int32),
        ("ns_var_noise_params", C.c_double * 8), ("ns_ls_noise_params", C.c_double * 8),
        ("use_aniso_mtx", C.c_int32), ("aniso_mtx", C.c_float * 9),
    ]


_vp = C.c_void_p


class MeanGrid(C.Structure):
    _fields_ = [("voxels", _vp), ("origin", C.c_int32 * 3), ("dims", C.c_int32 * 3),
                ("world_to_index", C.c_float * 12), ("clamp_min", C.c_float * 3), ("clamp_max", C.c_float * 3),
                ("offset", C.c_float), ("scale", C.c_float), ("volume", C.c_int32)]


class RaySoA(C.Structure):
    _fields_ = [(n, _vp) for n in (
        "ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "jitter", "noise_seed", "normal_seed",
        "flags", "prev_value", "prev_gx", "prev_gy", "prev_gz", "length_scale")]


class ReplySoA(C.Structure):
    _fields_ = [(n, _vp) for n in ("t", "value", "gx", "gy", "gz", "outcome", "evals")]


class PathReply(C.Structure):
    _fields_ = [(n, _vp) for n in ("t0", "gx0", "gy0", "gz0", "outcome0", "flights", "hits", "evals", "t_last",
                                   "outcome_last")]


class PathSeeds(C.Structure):
    _fields_ = [("px", _vp), ("py", _vp), ("spp", _vp), ("bounce", _vp), ("scene_seed", C.c_uint32)]


class Path1DReply(C.Structure):
    _fields_ = [(n, _vp) for n in ("t0", "gx0", "gy0", "gz0", "outcome0", "flights", "hits", "shadows", "visible",
                                   "nee", "t_last", "outcome_last", "evals")]


class NeeOut(C.Structure):
    _fields_ = [(n, _vp) for n in ("gx", "gy", "gz", "pdf")]


class NetCalibration(C.Structure):
    _fields_ = [("feat_mean", C.c_float * 8), ("feat_inv_std", C.c_float * 8), ("hull_lo", C.c_float * 8),
                ("hull_hi", C.c_float * 8), ("edges", C.c_float * 32), ("entropy_max", C.c_float)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "flights", "evals", "steps", "hits", "exits", "fails", "invalid", "kernel_launches", "spec_evals", "passes",
        "paired")]


_lib = None


class GpisError(RuntimeError):
    pass


def lib():
    """Load libgpis_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpisError(f"HIP extension missing: {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    st = C.c_int
    L.gpis_create.restype = st
    L.gpis_create.argtypes = [C.POINTER(FieldDesc), C.POINTER(_vp)]
    L.gpis_destroy.restype = st
    L.gpis_destroy.argtypes = [_vp]
    L.gpis_last_error.restype = C.c_char_p
    L.gpis_last_error.argtypes = [_vp]
    L.gpis_set_domain.restype = st
    L.gpis_set_domain.argtypes = [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32]
    L.gpis_impulse_table.restype = st
    L.gpis_impulse_table.argtypes = [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_uint32, _vp, _vp]
    L.gpis_eval_points.restype = st
    L.gpis_eval_points.argtypes = [_vp, _vp, _vp, _vp, C.c_int64, C.c_uint32, _vp, _vp, _vp, _vp, _vp]
    L.gpis_trace_flights.restype = st
    L.gpis_trace_flights.argtypes = [_vp, C.POINTER(RaySoA), C.c_int64, C.POINTER(ReplySoA), _vp]
    L.gpis_trace_paths.restype = st
    L.gpis_trace_paths.argtypes = [_vp, C.POINTER(RaySoA), C.POINTER(PathSeeds), C.c_int64, C.c_int32, C.c_uint32,
                                   C.c_float, C.POINTER(PathReply), _vp]
    L.gpis_trace_paths_1d.restype = st
    L.gpis_trace_paths_1d.argtypes = [_vp, C.POINTER(RaySoA), C.POINTER(PathSeeds), C.c_int64, C.c_int32,
                                      C.c_int32, C.c_uint32, C.c_float, C.POINTER(C.c_float),
                                      C.POINTER(Path1DReply), _vp]
    L.gpis_trace_flights_1d.restype = st
    L.gpis_trace_flights_1d.argtypes = [_vp, C.POINTER(RaySoA), C.c_int64, C.POINTER(ReplySoA),
                                        C.POINTER(C.c_float), C.POINTER(NeeOut), _vp]
    L.gpis_get_counters.restype = st
    L.gpis_get_counters.argtypes = [_vp, C.POINTER(Counters)]
    L.gpis_reset_counters.restype = st
    L.gpis_reset_counters.argtypes = [_vp]
    L.gpis_set_launch.restype = st
    L.gpis_set_launch.argtypes = [_vp, C.c_int32, C.c_int32]
    L.gpis_set_refill.restype = st
    L.gpis_set_refill.argtypes = [_vp, C.c_int32]
    L.gpis_set_coop_cells.restype = st
    L.gpis_set_coop_cells.argtypes = [_vp, C.c_int32]
    L.gpis_set_lists.restype = st
    L.gpis_set_lists.argtypes = [_vp, C.c_int32, C.c_int32]
    L.gpis_set_precision.restype = st
    L.gpis_set_precision.argtypes = [_vp, C.c_int32]
    L.gpis_set_mean_grid.restype = st
    L.gpis_set_mean_grid.argtypes = [_vp, C.POINTER(MeanGrid)]
    L.gpis_net_create.restype = st
    L.gpis_net_create.argtypes = [C.c_uint64, C.POINTER(_vp)]
    L.gpis_net_destroy.restype = st
    L.gpis_net_destroy.argtypes = [_vp]
    L.gpis_net_last_error.restype = C.c_char_p
    L.gpis_net_last_error.argtypes = [_vp]
    L.gpis_net_set_params.restype = st
    L.gpis_net_set_params.argtypes = [_vp, C.POINTER(C.c_float), C.c_int64]
    L.gpis_net_get_params.restype = st
    L.gpis_net_get_params.argtypes = [_vp, C.POINTER(C.c_float), C.c_int64]
    L.gpis_net_set_calibration.restype = st
    L.gpis_net_set_calibration.argtypes = [_vp, C.POINTER(NetCalibration)]
    L.gpis_net_features_1d.restype = st
    L.gpis_net_features_1d.argtypes = [_vp, C.POINTER(RaySoA), C.c_int64, _vp, _vp]
    L.gpis_net_classes_1d.restype = st
    L.gpis_net_classes_1d.argtypes = [_vp, C.POINTER(RaySoA), C.POINTER(ReplySoA), C.c_int64, _vp, _vp]
    L.gpis_net_gradient.restype = st
    L.gpis_net_gradient.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp]
    L.gpis_net_train_step.restype = st
    L.gpis_net_train_step.argtypes = [_vp, _vp, _vp, C.c_int64, C.c_float, _vp, _vp]
    L.gpis_net_set_averaging.restype = st
    L.gpis_net_set_averaging.argtypes = [_vp, C.c_float]
    L.gpis_net_use_average.restype = st
    L.gpis_net_use_average.argtypes = [_vp]
    L.gpis_net_forward.restype = st
    L.gpis_net_forward.argtypes = [_vp, _vp, C.c_int64, _vp, _vp]
    L.gpis_surrogate_flights_1d.restype = st
    L.gpis_surrogate_flights_1d.argtypes = [_vp, _vp, C.POINTER(RaySoA), C.c_int64, C.POINTER(ReplySoA),
                                            C.POINTER(C.c_uint32), _vp]
    L.gpis_surrogate_flights_1d_noalloc.restype = st
    L.gpis_surrogate_flights_1d_noalloc.argtypes = [_vp, _vp, C.POINTER(RaySoA), C.c_int64, C.POINTER(ReplySoA),
                                            C.POINTER(C.c_uint32), _vp]
    _lib = L
    return L


def check(status: int, ctx=None, what: str = "", net=None) -> None:
    if status != GPIS_OK:
        msg = ""
        if net:
            msg = lib().gpis_net_last_error(net).decode(errors="replace")
        if ctx and not msg:
            msg = lib().gpis_last_error(ctx).decode(errors="replace")
        raise GpisError(f"{what}: {STATUS_NAMES.get(status, status)} {msg}".strip())
