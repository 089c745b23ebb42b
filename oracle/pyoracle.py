"""ctypes bindings for the CPU oracle (oracle/build/libgpis_oracle.so) and the
reference leaf library (oracle/_ref/libgpis_ref_leaf.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the gpis_extend_amd product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libgpis_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgpis_ref_leaf.so")

COV_SE, COV_PROC_NS = 0, 1
MEAN_SPHERE, MEAN_BOX, MEAN_PLANE, MEAN_GRID, MEAN_PROC = 0, 1, 2, 3, 4
# SdfFunctions::Function (SdfFunctions.hpp:13-22)
(SDF_KNOB, SDF_KNOB_INNER, SDF_KNOB_OUTER, SDF_TWO_SPHERES, SDF_PLANE, SDF_ROUNDED_U_GROOVE, SDF_TWO_SLABS,
 SDF_THREE_SLABS) = range(8)
CTX_GLOBAL, CTX_RENEWAL, CTX_RENEWAL_PLUS, CTX_NONE = 0, 1, 2, 3
SCHEME_UNI, SCHEME_NEE, SCHEME_MIS = 0, 1, 2
LS_CONST, LS_SINCOS, LS_NOISE = 0, 1, 2
VAR_CONST, VAR_NOISE = 0, 1
NOISE_TYPES = ("bottom_top", "left_right", "front_back", "bottom_top_left_right", "sandstone", "rust")
EXIT, HIT, FAIL = 0, 1, 2
FAIL_NONE, FAIL_OUTWARD, FAIL_ZERO_GRAD, FAIL_NONFINITE, FAIL_BOUNCE_LIMIT = 0, 1, 2, 3, 4


class MediumDesc(C.Structure):
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
        ("grid", C.c_void_p), ("grid_origin", C.c_int32 * 3), ("grid_dims", C.c_int32 * 3),
        ("grid_w2i", C.c_float * 12), ("grid_clamp_min", C.c_float * 3), ("grid_clamp_max", C.c_float * 3),
        ("grid_offset", C.c_float), ("grid_scale", C.c_float), ("grid_volume", C.c_int32),
        ("iso3d", C.c_int32),
        ("sdf_function", C.c_int32), ("sdf_inverse_transform", C.c_float * 12),
        ("sdf_min", C.c_float), ("sdf_scale", C.c_float), ("sdf_offset", C.c_float),
        ("ns_var_field", C.c_int32), ("ns_var_noise", C.c_int32), ("ns_ls_noise", C.c_int32),
        ("ns_var_noise_params", C.c_double * 8), ("ns_ls_noise_params", C.c_double * 8),
        ("use_aniso_mtx", C.c_int32), ("aniso_mtx", C.c_float * 9),
    ]


class Ray(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("dir", C.c_float * 3),
        ("near_t", C.c_float), ("far_t", C.c_float), ("jitter", C.c_float),
        ("pss", C.c_uint32 * 4), ("scene_seed", C.c_uint32), ("first_scatter", C.c_int32),
        ("last_val", C.c_float), ("last_aniso", C.c_double * 3), ("bounce", C.c_int32),
    ]


class Flight(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("value", C.c_float), ("grad", C.c_float * 3),
        ("outcome", C.c_int32), ("fail_reason", C.c_int32), ("steps", C.c_int32),
        ("evals", C.c_int32), ("scheme", C.c_int32), ("min_abs_f", C.c_float),
        ("bracket", C.c_float * 2),
    ]


RAY_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("dir", "<f4", 3), ("near_t", "<f4"), ("far_t", "<f4"), ("jitter", "<f4"),
    ("pss", "<u4", 4), ("scene_seed", "<u4"), ("first_scatter", "<i4"), ("last_val", "<f4"),
    ("last_aniso", "<f8", 3), ("bounce", "<i4")], align=True)
FLIGHT_DTYPE = np.dtype([
    ("t", "<f8"), ("value", "<f4"), ("grad", "<f4", 3), ("outcome", "<i4"), ("fail_reason", "<i4"),
    ("steps", "<i4"), ("evals", "<i4"), ("scheme", "<i4"), ("min_abs_f", "<f4"),
    ("bracket", "<f4", 2)], align=True)
PATH_DTYPE = np.dtype([
    ("t0", "<f4"), ("grad0", "<f4", 3), ("outcome0", "<i4"), ("fail0", "<i4"),
    ("flights", "<i4"), ("hits", "<i4"), ("shadows", "<i4"), ("visible", "<i4"),
    ("nee", "<f4"), ("t_last", "<f4"), ("outcome_last", "<i4"), ("fail_last", "<i4"),
    ("evals", "<i4"), ("pos_last", "<f4", 3), ("dir_last", "<f4", 3)], align=True)
PHASE_MIRROR, PHASE_LAMBERT = 0, 1
assert RAY_DTYPE.itemsize == C.sizeof(Ray), (RAY_DTYPE.itemsize, C.sizeof(Ray))
assert FLIGHT_DTYPE.itemsize == C.sizeof(Flight), (FLIGHT_DTYPE.itemsize, C.sizeof(Flight))

_lib = None
_ref = None


def use_library(path):
    """Load the oracle from `path` (another build of the same source) instead of the default
    build; only before the first call into the oracle."""
    global LIB_PATH
    if _lib is not None and os.path.abspath(path) != os.path.abspath(LIB_PATH):
        raise RuntimeError("the oracle library is already loaded")
    LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        u32, f32p, vp = C.c_uint32, C.POINTER(C.c_float), C.c_void_p
        L.orc_xxhash32_1.restype = u32; L.orc_xxhash32_1.argtypes = [u32]
        L.orc_xxhash32_2.restype = u32; L.orc_xxhash32_2.argtypes = [u32, u32]
        L.orc_xxhash32_4.restype = u32; L.orc_xxhash32_4.argtypes = [u32] * 4
        L.orc_cell3d_impulses.argtypes = [u32, u32, u32, u32, u32, f32p]
        L.orc_cell1d_impulses.argtypes = [u32, u32, u32, f32p]
        L.orc_standard_normals.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double)]
        L.orc_transverse_moments.restype = C.c_int
        L.orc_transverse_moments.argtypes = [vp, vp, f32p, vp, vp, vp]
        L.orc_grid_density.restype = C.c_float
        L.orc_grid_density.argtypes = [vp, f32p]
        L.orc_grid_mean.restype = C.c_double
        L.orc_grid_mean.argtypes = [vp, vp]
        L.orc_grid_mean_grad.argtypes = [vp, vp, vp]
        L.orc_tangent_frame.argtypes = [f32p, f32p]
        L.orc_field_create.restype = vp; L.orc_field_create.argtypes = [C.POINTER(MediumDesc)]
        L.orc_field_destroy.argtypes = [vp]
        L.orc_compute_seed.restype = u32
        L.orc_compute_seed.argtypes = [vp, C.POINTER(C.c_uint32), u32]
        L.orc_noise3d.argtypes = [vp, f32p, u32, f32p]
        L.orc_eval_points.argtypes = [vp, f32p, C.c_int64, u32, f32p, C.c_int]
        L.orc_eval_points_dir.argtypes = [vp, f32p, C.c_int64, u32, f32p, f32p, C.c_int]
        L.orc_kernel_radius.restype = C.c_float; L.orc_kernel_radius.argtypes = [vp]
        L.orc_trace_flight.argtypes = [vp, C.POINTER(Ray), C.POINTER(Flight)]
        L.orc_trace_flights.argtypes = [vp, vp, C.c_int64, vp, C.c_int]
        L.orc_trace_flights_ls.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int]
        L.orc_trace_paths1d.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, u32, C.c_float, vp, vp,
                                        C.c_int]
        L.orc_nee_ls.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, vp, vp]
        L.orc_trace_paths3d.argtypes = [vp, vp, C.c_int64, C.c_int32, u32, C.c_float, C.c_int32, C.c_float, u32, vp,
                                        C.c_int]
        L.orc_sincos_turn.argtypes = [C.c_float, f32p, f32p]
        L.orc_diffuse_direction.argtypes = [u32, u32, u32, f32p, f32p, f32p]
        L.orc_conditioned_eval_ls.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp]
        L.orc_trace_flight_perturbed.argtypes = [vp, C.POINTER(Ray), C.c_float, C.c_uint32, C.POINTER(Flight)]
        L.orc_conditioned_eval.argtypes = [vp, C.POINTER(Ray), f32p, C.c_float, f32p]
        L.orc_nee.argtypes = [vp, C.POINTER(Ray), f32p, f32p, C.c_float, f32p, f32p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        u32, f32p = C.c_uint32, C.POINTER(C.c_float)
        L.ref_xxhash32_1.restype = u32; L.ref_xxhash32_1.argtypes = [u32]
        L.ref_xxhash32_2.restype = u32; L.ref_xxhash32_2.argtypes = [u32, u32]
        L.ref_xxhash32_4.restype = u32; L.ref_xxhash32_4.argtypes = [u32] * 4
        L.ref_cell3d_impulses.argtypes = [u32, u32, u32, u32, u32, f32p]
        L.ref_cell1d_impulses.argtypes = [u32, u32, u32, f32p]
        L.ref_standard_normals.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double)]
        L.ref_tangent_frame.argtypes = [f32p, f32p]
        L.ref_pcg_words.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_uint32)]
        L.ref_random3.restype = L.ref_simplex3d.restype = L.ref_fbm.restype = None
        _ref = L
    return _ref


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def cell3d(i, j, k, seed, n, which="oracle"):
    out = np.zeros(4 * n, np.float32)
    L = lib() if which == "oracle" else ref()
    fn = L.orc_cell3d_impulses if which == "oracle" else L.ref_cell3d_impulses
    fn(i & 0xFFFFFFFF, j & 0xFFFFFFFF, k & 0xFFFFFFFF, seed & 0xFFFFFFFF, n, _fp(out))
    return out.reshape(n, 4)


def cell1d(i, seed, n, which="oracle"):
    out = np.zeros(2 * n, np.float32)
    L = lib() if which == "oracle" else ref()
    fn = L.orc_cell1d_impulses if which == "oracle" else L.ref_cell1d_impulses
    fn(i & 0xFFFFFFFF, seed & 0xFFFFFFFF, n, _fp(out))
    return out.reshape(n, 2)


def standard_normals(state, npairs, which="oracle"):
    out = np.zeros(2 * npairs, np.float64)
    L = lib() if which == "oracle" else ref()
    fn = L.orc_standard_normals if which == "oracle" else L.ref_standard_normals
    fn(state, npairs, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def tangent_frame(n, which="oracle"):
    nn = np.asarray(n, np.float32).copy()
    out = np.zeros(9, np.float32)
    L = lib() if which == "oracle" else ref()
    fn = L.orc_tangent_frame if which == "oracle" else L.ref_tangent_frame
    fn(_fp(nn), _fp(out))
    return out.reshape(3, 3)


def make_desc(**kw) -> MediumDesc:
    """Medium description with the reference defaults (SparseConvolutionNoiseMedium.cpp:62-79,
    SquaredExponentialCovariance defaults, localScale 3)."""
    d = MediumDesc()
    d.cov_kind = kw.get("cov_kind", COV_SE)
    d.sigma = kw.get("sigma", 1.0)
    d.length_scale = kw.get("length_scale", 1.0)
    d.kernel_scale = kw.get("kernel_scale", 3.0)
    for i, a in enumerate(kw.get("aniso", (1.0, 1.0, 1.0))):
        d.aniso[i] = a
    d.ns_multires = kw.get("ns_multires", 0)
    d.ns_ls_field = kw.get("ns_ls_field", LS_CONST)
    for i, a in enumerate(kw.get("ns_ls_params", (1.0, 0.0, 0.0, 0.0))):
        d.ns_ls_params[i] = a
    d.ns_ls_max = kw.get("ns_ls_max", 1.0)
    d.ns_var = kw.get("ns_var", 1.0)
    d.mean_kind = kw.get("mean_kind", MEAN_SPHERE)
    d.mean_radius = kw.get("mean_radius", 1.0)
    for i, a in enumerate(kw.get("mean_center", (0.0, 0.0, 0.0))):
        d.mean_center[i] = a
    for i, a in enumerate(kw.get("mean_half", (0.5, 0.5, 0.5))):
        d.mean_half[i] = a
    for i, a in enumerate(kw.get("plane_normal", (1.0, 0.0, 0.0))):
        d.plane_normal[i] = a
    d.plane_offset = kw.get("plane_offset", 0.0)
    d.global_seed = kw.get("global_seed", 0)
    d.impulse_density = kw.get("impulse_density", 3.0)
    d.single_realization = kw.get("single_realization", 0)
    d.ctxt = kw.get("ctxt", CTX_RENEWAL_PLUS)
    d.one_d = kw.get("one_d", 0)
    d.scheme = kw.get("scheme", SCHEME_UNI)
    d.correlation_xy = kw.get("correlation_xy", 0)
    d.step_size = kw.get("step_size", 0.01)
    d.min_step = kw.get("min_step", 8)
    d.max_bounce = kw.get("max_bounce", 1 << 30)
    d.iso3d = int(kw.get("iso3d", 0))
    d.sdf_function = int(kw.get("sdf_function", SDF_KNOB))
    for i, a in enumerate(kw.get("sdf_inverse_transform", (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0))):
        d.sdf_inverse_transform[i] = a
    d.sdf_min = kw.get("sdf_min", -3.4028234663852886e38)
    d.sdf_scale = kw.get("sdf_scale", 1.0)
    d.sdf_offset = kw.get("sdf_offset", 0.0)
    d.ns_var_field = kw.get("ns_var_field", VAR_CONST)
    d.ns_var_noise = kw.get("ns_var_noise", 0)
    d.ns_ls_noise = kw.get("ns_ls_noise", 0)
    for i, a in enumerate(kw.get("ns_var_noise_params", (1.0, 500.0, 0.0, 1.0, 1.0, 500.0, 0.0, 1.0))):
        d.ns_var_noise_params[i] = a
    for i, a in enumerate(kw.get("ns_ls_noise_params", (1.0, 500.0, 0.0, 1.0, 1.0, 500.0, 0.0, 1.0))):
        d.ns_ls_noise_params[i] = a
    m = kw.get("aniso_mtx")
    d.use_aniso_mtx = int(m is not None)
    for i, a in enumerate(np.asarray(m if m is not None else np.eye(3), np.float64).reshape(9)):
        d.aniso_mtx[i] = float(a)
    g = kw.get("grid")
    if g is not None:
        set_grid(d, g.voxels, g.origin, g.world_to_index, g.clamp_min, g.clamp_max,
                 kw.get("grid_offset", 0.0), kw.get("grid_scale", 1.0), kw.get("grid_volume", False))
    return d


def set_grid(d: MediumDesc, voxels, origin, world_to_index, clamp_min, clamp_max, offset=0.0, scale=1.0,
             volume=False):
    """Attach a dense tabulated-mean grid (voxels indexed [z][y][x]) to a description;
    the description keeps the array alive."""
    v = np.ascontiguousarray(voxels, np.float32)
    assert v.ndim == 3
    d._grid_keep = v
    d.grid = v.ctypes.data
    for i in range(3):
        d.grid_origin[i] = int(origin[i])
        d.grid_dims[i] = v.shape[2 - i]
        d.grid_clamp_min[i] = clamp_min[i]
        d.grid_clamp_max[i] = clamp_max[i]
    w = np.asarray(world_to_index, np.float32).reshape(12)
    for i in range(12):
        d.grid_w2i[i] = w[i]
    d.grid_offset, d.grid_scale, d.grid_volume = offset, scale, int(bool(volume))
    return d


def grid_density(d: MediumDesc, p) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    return np.array([lib().orc_grid_density(C.byref(d), _fp(p[i].copy())) for i in range(len(p))], np.float32)


def grid_mean(d: MediumDesc, a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 3)
    return np.array([lib().orc_grid_mean(C.byref(d), a[i].ctypes.data) for i in range(len(a))])


def grid_mean_grad(d: MediumDesc, a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 3)
    g = np.zeros((len(a), 3))
    for i in range(len(a)):
        row = np.zeros(3)
        lib().orc_grid_mean_grad(C.byref(d), a[i].ctypes.data, row.ctypes.data)
        g[i] = row
    return g


class Field:
    def __init__(self, desc: MediumDesc):
        self.desc = desc
        self.h = lib().orc_field_create(C.byref(desc))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_field_destroy(self.h)
            self.h = None

    @property
    def kernel_radius(self) -> float:
        return lib().orc_kernel_radius(self.h)

    def compute_seed(self, pss, scene_seed):
        a = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in pss])
        return lib().orc_compute_seed(self.h, a, scene_seed & 0xFFFFFFFF)

    def noise3d(self, p, seed):
        pp = np.asarray(p, np.float32).copy()
        out = np.zeros(4, np.float32)
        lib().orc_noise3d(self.h, _fp(pp), seed, _fp(out))
        return out

    def eval_points(self, xyz, seed, nthreads=1):
        xyz = np.ascontiguousarray(xyz, np.float32)
        out = np.zeros((xyz.shape[0], 4), np.float32)
        lib().orc_eval_points(self.h, _fp(xyz), xyz.shape[0], seed, _fp(out), nthreads)
        return out

    def eval_points_dir(self, xyz, seed, direction, nthreads=1):
        """Value and gradient at points for flights along `direction` (the isotropic
        ray space of isotropic_3D_sampling media depends on it)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        d = np.asarray(direction, np.float32).copy()
        out = np.zeros((xyz.shape[0], 4), np.float32)
        lib().orc_eval_points_dir(self.h, _fp(xyz), xyz.shape[0], seed, _fp(d), _fp(out), nthreads)
        return out

    def trace(self, rays: np.ndarray, nthreads=1) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.shape[0], FLIGHT_DTYPE)
        lib().orc_trace_flights(self.h, rays.ctypes.data, rays.shape[0], out.ctypes.data, nthreads)
        return out

    def trace_paths3d(self, rays: np.ndarray, max_bounces, bounce_seed, box, seeded=False, nthreads=1,
                      perturb=0.0, pseed=0) -> np.ndarray:
        """3D paths (orc_trace_paths3d): path i's bounces use path index i.  `perturb`:
        +-perturb added to every marched value (signs hashed from pseed, path, flight)."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.shape[0], PATH_DTYPE)
        lib().orc_trace_paths3d(self.h, rays.ctypes.data, rays.shape[0], int(max_bounces),
                                int(bounce_seed) & 0xFFFFFFFF, float(box), int(bool(seeded)), float(perturb),
                                int(pseed) & 0xFFFFFFFF, out.ctypes.data, nthreads)
        return out

    def trace_perturbed(self, ray_rec, eps, pseed):
        r = Ray.from_buffer_copy(np.ascontiguousarray(ray_rec, RAY_DTYPE).tobytes())
        out = np.zeros(1, FLIGHT_DTYPE)
        lib().orc_trace_flight_perturbed(self.h, C.byref(r), eps, pseed, out.ctypes.data_as(C.POINTER(Flight)))
        return out[0]

    def conditioned_eval(self, ray_rec, q, t_segment=1.0):
        r = Ray.from_buffer_copy(np.ascontiguousarray(ray_rec, RAY_DTYPE).tobytes())
        qq = np.asarray(q, np.float32).copy()
        out = np.zeros(4, np.float32)
        lib().orc_conditioned_eval(self.h, C.byref(r), _fp(qq), t_segment, _fp(out))
        return out

    def transverse_moments(self, ray_rec, p):
        """(applies, mean[2], variance, sampling scheme) of the correlated transverse
        gradient law at p for the ray's conditioned realization."""
        r = Ray.from_buffer_copy(np.ascontiguousarray(ray_rec, RAY_DTYPE).tobytes())
        pp = np.asarray(p, np.float32).copy()
        mu = np.zeros(2)
        var = np.zeros(1)
        sch = C.c_int(0)
        ok = lib().orc_transverse_moments(self.h, C.byref(r), _fp(pp), mu.ctypes.data, var.ctypes.data,
                                          C.byref(sch))
        return bool(ok), mu, float(var[0]), sch.value

    def nee(self, ray_rec, normal, p, t_segment):
        r = Ray.from_buffer_copy(np.ascontiguousarray(ray_rec, RAY_DTYPE).tobytes())
        nn = np.asarray(normal, np.float32).copy()
        pp = np.asarray(p, np.float32).copy()
        g = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        lib().orc_nee(self.h, C.byref(r), _fp(nn), _fp(pp), t_segment, _fp(g), _fp(pdf))
        return g, float(pdf[0])


def trace_ls(desc: MediumDesc, rays: np.ndarray, ls, nthreads=1) -> np.ndarray:
    """Flights with a per-ray SE length scale `ls` (None: the medium's own)."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = np.zeros(rays.shape[0], FLIGHT_DTYPE)
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float32)
    lib().orc_trace_flights_ls(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data,
                               rays.shape[0], out.ctypes.data, nthreads)
    return out


def trace_paths1d(desc: MediumDesc, rays: np.ndarray, ls, max_bounces, phase, bounce_seed, box, light=None,
                  nthreads=1) -> np.ndarray:
    """1D paths (orc_trace_paths1d): path i's draws use path index i."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = np.zeros(rays.shape[0], PATH_DTYPE)
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float32)
    la = None if light is None else np.ascontiguousarray(light, np.float32)
    lib().orc_trace_paths1d(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data,
                            rays.shape[0], int(max_bounces), int(phase), int(bounce_seed) & 0xFFFFFFFF, float(box),
                            None if la is None else la.ctypes.data, out.ctypes.data, nthreads)
    return out


def sincos_turn(u: float):
    """(sin, cos) of 2 pi u by the bounce's fixed polynomial."""
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_turn(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def diffuse_direction(ray, bounce, bounce_seed, normal):
    """(direction, jitter) of the paths' diffuse bounce about a unit normal."""
    n = np.ascontiguousarray(normal, np.float32)
    d = np.zeros(3, np.float32)
    j = C.c_float()
    lib().orc_diffuse_direction(int(ray) & 0xFFFFFFFF, int(bounce) & 0xFFFFFFFF, int(bounce_seed) & 0xFFFFFFFF,
                                _fp(n), _fp(d), C.byref(j))
    return d, j.value


def nee_ls(desc: MediumDesc, rays: np.ndarray, ls, normals, points, t_segment):
    """neeGrad / neePDF per ray with a per-ray SE length scale."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = rays.shape[0]
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float32)
    nn = np.ascontiguousarray(normals, np.float32).reshape(n, 3)
    pp = np.ascontiguousarray(points, np.float32).reshape(n, 3)
    ts = np.ascontiguousarray(t_segment, np.float32).reshape(n)
    g = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_nee_ls(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data, nn.ctypes.data,
                     pp.ctypes.data, ts.ctypes.data, n, g.ctypes.data, pdf.ctypes.data)
    return g, pdf


def conditioned_eval_ls(desc: MediumDesc, rays: np.ndarray, ls, q, t_segment):
    """Value and gradient of each ray's conditioned realization at q[i] (per-ray l)."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = rays.shape[0]
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float32)
    qq = np.ascontiguousarray(q, np.float32).reshape(n, 3)
    ts = np.ascontiguousarray(t_segment, np.float32).reshape(n)
    out = np.zeros((n, 4), np.float32)
    lib().orc_conditioned_eval_ls(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data,
                                  qq.ctypes.data, ts.ctypes.data, n, out.ctypes.data)
    return out


# ---- FP64 truth build (gpis_oracle.c with ORC_FP64_TRUTH) ----------------
LIB64_PATH = os.path.join(HERE, "build", "libgpis_oracle_fp64.so")
_lib64 = None


def lib64():
    """The FP64 truth build: the same algorithm with every float of the field arithmetic
    in double; the request/reply structs are those of the float build, the pointer
    arguments that the float build takes as float are double here."""
    global _lib64
    if _lib64 is None:
        if not os.path.exists(LIB64_PATH):
            raise RuntimeError(f"FP64 oracle not built: {LIB64_PATH} (run make -C oracle)")
        L = C.CDLL(LIB64_PATH)
        vp, u32 = C.c_void_p, C.c_uint32
        L.orc_field_create.restype = vp
        L.orc_field_create.argtypes = [C.POINTER(MediumDesc)]
        L.orc_field_destroy.argtypes = [vp]
        L.orc_eval_points.argtypes = [vp, vp, C.c_int64, u32, vp, C.c_int]
        L.orc_trace_flights.argtypes = [vp, vp, C.c_int64, vp, C.c_int]
        L.orc_trace_flights_ls.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int]
        L.orc_conditioned_eval_ls.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp]
        _lib64 = L
    return _lib64


class Truth64:
    """A field of the FP64 truth build: value and gradient in double, flights whose
    distance t is the double march's (value and gradient of the reply rounded to float)."""

    def __init__(self, desc: MediumDesc):
        self.desc = desc
        self.h = lib64().orc_field_create(C.byref(desc))

    def __del__(self):
        if getattr(self, "h", None):
            lib64().orc_field_destroy(self.h)
            self.h = None

    def eval_points(self, xyz, seed, nthreads=1):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 4), np.float64)
        lib64().orc_eval_points(self.h, xyz.ctypes.data, xyz.shape[0], int(seed) & 0xFFFFFFFF, out.ctypes.data,
                                nthreads)
        return out

    def trace(self, rays: np.ndarray, nthreads=1) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.shape[0], FLIGHT_DTYPE)
        lib64().orc_trace_flights(self.h, rays.ctypes.data, rays.shape[0], out.ctypes.data, nthreads)
        return out


def trace_ls64(desc: MediumDesc, rays: np.ndarray, ls, nthreads=1) -> np.ndarray:
    """trace_ls in the FP64 truth build."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = np.zeros(rays.shape[0], FLIGHT_DTYPE)
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float64)
    lib64().orc_trace_flights_ls(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data,
                                 rays.shape[0], out.ctypes.data, nthreads)
    return out


def conditioned_eval_ls64(desc: MediumDesc, rays: np.ndarray, ls, q, t_segment):
    """conditioned_eval_ls in the FP64 truth build (double in and out)."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = rays.shape[0]
    lsa = None if ls is None else np.ascontiguousarray(ls, np.float64)
    qq = np.ascontiguousarray(q, np.float64).reshape(n, 3)
    ts = np.ascontiguousarray(t_segment, np.float64).reshape(n)
    out = np.zeros((n, 4), np.float64)
    lib64().orc_conditioned_eval_ls(C.byref(desc), rays.ctypes.data, None if lsa is None else lsa.ctypes.data,
                                    qq.ctypes.data, ts.ctypes.data, n, out.ctypes.data)
    return out


# ---- ProceduralNoise fields (oracle restatement, and the reference build's fbm) ----
def _noise_io(p):
    return np.ascontiguousarray(p, np.float64).reshape(-1, 3)


def random3(p, which="oracle"):
    L = lib() if which == "oracle" else ref()
    q = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    out = np.zeros_like(q)
    fn = L.orc_random3 if which == "oracle" else L.ref_random3
    fn(q.ctypes.data_as(C.POINTER(C.c_float)), C.c_int64(len(q)), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def simplex3d(p, which="oracle"):
    L = lib() if which == "oracle" else ref()
    q = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    out = np.zeros(len(q), np.float32)
    fn = L.orc_simplex3d if which == "oracle" else L.ref_simplex3d
    fn(q.ctypes.data_as(C.POINTER(C.c_float)), C.c_int64(len(q)), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def fbm(p, octaves, which="oracle"):
    L = lib() if which == "oracle" else ref()
    q = _noise_io(p)
    out = np.zeros(len(q), np.float64)
    fn = L.orc_fbm if which == "oracle" else L.ref_fbm
    fn(q.ctypes.data_as(C.POINTER(C.c_double)), C.c_int64(len(q)), C.c_int(octaves),
       out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def noise_field(kind, params8, p, vec=False):
    """ProceduralNoise (vec False) or the largest component of ProceduralNoiseVec (vec True)
    of type `kind` (an index of NOISE_TYPES) at points p."""
    q = _noise_io(p)
    out = np.zeros(len(q), np.float64)
    prm = (C.c_double * 8)(*[float(v) for v in params8])
    lib().orc_noise_field(C.c_int32(kind), prm, C.c_int(1 if vec else 0), q.ctypes.data_as(C.POINTER(C.c_double)),
                          C.c_int64(len(q)), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out
