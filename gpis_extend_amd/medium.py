"""Host-side mirror of Tungsten's ``sparse_conv_noise`` medium over the HIP C ABI.

``SparseConvolutionNoiseMedium.from_json`` reads the same JSON parameters as the
reference (src/core/media/SparseConvolutionNoiseMedium.cpp:102-118,
src/core/media/GaussianProcessMedium.cpp:98-127, src/core/math/GPFunctions.cpp:657-685,
src/core/math/GaussianProcess.cpp:173-191) and fails on the same invalid values.
``GpisDevice`` owns one gpis_ctx and exposes the wavefront entry points on torch
device tensors: ``sample_distance`` (GaussianProcessMedium::sampleDistance, :222-234)
and ``transmittance`` (:424-440) semantics, batched.

torch is used only for device memory and the current HIP stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import _capi as A
from . import mat4f
from .hashing import xxhash32_1, xxhash32_4


class MediumConfigError(ValueError):
    pass


_CTX = {"global": A.CTX_GLOBAL, "renewal+": A.CTX_RENEWAL_PLUS, "renewal": A.CTX_RENEWAL, "none": A.CTX_NONE}
_SCHEME = {"uni": A.SCHEME_UNI, "UNI": A.SCHEME_UNI, "nee": A.SCHEME_NEE, "NEE": A.SCHEME_NEE,
           "mis": A.SCHEME_MIS, "MIS": A.SCHEME_MIS}


@dataclass
class DenseMeanGrid:
    """A tabulated mean's voxel grid in the dense layout of the reference's GPU export
    (HostMeanGrid, src/gpu/cuda/RaySegmentBatch.hpp:19-30, filled by makeHostMeanGrid,
    src/gpu/VdbMeanGrid.cpp:9-32): ``voxels[z][y][x]`` hold the density at integer index
    coordinates ``origin + (x, y, z)``; a world point p is looked up at the index point
    ``world_to_index @ (p, 1)`` clamped to ``[clamp_min, clamp_max]``."""
    voxels: np.ndarray
    origin: tuple = (0, 0, 0)
    world_to_index: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    clamp_min: tuple = (0.0, 0.0, 0.0)
    clamp_max: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        self.voxels = np.ascontiguousarray(self.voxels, np.float32)
        if self.voxels.ndim != 3 or min(self.voxels.shape) < 2:
            raise MediumConfigError("a dense mean grid needs at least 2 voxels along each axis")
        self.origin = tuple(int(v) for v in self.origin)
        w = np.asarray(self.world_to_index, np.float64).reshape(-1)
        if w.size == 16:
            w = w[:12]
        if w.size != 12 or not np.all(np.isfinite(w)):
            raise MediumConfigError("world_to_index must be a finite 3x4 (or 4x4) row-major matrix")
        self.world_to_index = tuple(float(v) for v in w)
        self.clamp_min = tuple(float(v) for v in self.clamp_min)
        self.clamp_max = tuple(float(v) for v in self.clamp_max)
        if any(a > b for a, b in zip(self.clamp_min, self.clamp_max)):
            raise MediumConfigError("empty clamp box")

    @property
    def dims(self) -> tuple:
        return (self.voxels.shape[2], self.voxels.shape[1], self.voxels.shape[0])

    @classmethod
    def sample(cls, fn, lo, hi, spacing: float) -> "DenseMeanGrid":
        """Tabulate fn(x, y, z) (vectorised, world units) on the lattice lo + spacing * i
        covering [lo, hi]; lookups clamp to the lattice."""
        lo = np.asarray(lo, np.float64)
        n = np.floor((np.asarray(hi, np.float64) - lo) / spacing + 1e-9).astype(int) + 1
        axes = [lo[i] + spacing * np.arange(n[i]) for i in range(3)]
        z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        vox = np.asarray(fn(x, y, z), np.float32)
        inv = 1.0 / spacing
        w2i = (inv, 0.0, 0.0, -lo[0] * inv, 0.0, inv, 0.0, -lo[1] * inv, 0.0, 0.0, inv, -lo[2] * inv)
        return cls(vox, (0, 0, 0), w2i, (0.0, 0.0, 0.0), tuple(float(v - 1) for v in n))

    @classmethod
    def from_json(cls, js: dict, base_dir: Optional[str] = None) -> "DenseMeanGrid":
        """``{"type": "dense", "file": "<voxels .npy, [z][y][x]>", "origin": [...],
        "world_to_index": [12 numbers], "clamp_min": [...], "clamp_max": [...]}``."""
        import os
        path = js["file"]
        if base_dir and not os.path.isabs(path):
            path = os.path.join(base_dir, path)
        vox = np.load(path, allow_pickle=False)
        shape = vox.shape
        return cls(vox, tuple(js.get("origin", (0, 0, 0))),
                   tuple(js.get("world_to_index", cls.world_to_index)),
                   tuple(js.get("clamp_min", (0.0, 0.0, 0.0))),
                   tuple(js.get("clamp_max", (shape[2] - 1.0, shape[1] - 1.0, shape[0] - 1.0))))

    def to_c(self, offset: float, scale: float, volume: bool) -> A.MeanGrid:
        g = A.MeanGrid()
        g.voxels = self.voxels.ctypes.data
        for i in range(3):
            g.origin[i], g.dims[i] = self.origin[i], self.dims[i]
            g.clamp_min[i], g.clamp_max[i] = self.clamp_min[i], self.clamp_max[i]
        for i in range(12):
            g.world_to_index[i] = self.world_to_index[i]
        g.offset, g.scale, g.volume = offset, scale, int(bool(volume))
        return g


def _noise_field(js):
    """A ProceduralNoise / ProceduralNoiseVec of type "noise" (GPFunctions.hpp:716-735): its
    type index and (min, max, start, end, min2, max2, start2, end2) with the reference's
    defaults."""
    kind = js.get("noise", "bottom_top")
    if kind not in A.NOISE_TYPES:
        raise MediumConfigError(f"Invalid noise typ function: '{kind}'")
    prm = tuple(float(js.get(k, d)) for k, d in (("min", 1.0), ("max", 500.0), ("start", 0.0), ("end", 1.0),
                                                ("min2", 1.0), ("max2", 500.0), ("start2", 0.0), ("end2", 1.0)))
    return A.NOISE_TYPES.index(kind), prm


def _noise_max(kind, prm):
    """ProceduralNoiseVec::maxVal() (GPFunctions.cpp:121-139), as a float."""
    if kind <= 2:
        return float(np.float32(max(prm[1], prm[0])))
    if kind == 3:
        return float(np.float32(max(prm[1], prm[0]) * max(prm[5], prm[4])))
    return 1.0


@dataclass
class SparseConvolutionNoiseMedium:
    """Parameters of one sparse_conv_noise medium.  Defaults are the reference
    constructor's (SparseConvolutionNoiseMedium.cpp:62-79, Medium.cpp:24,
    GaussianProcessMedium.cpp:87-96)."""
    step_size: float = 0.01
    min_step: int = 8
    seed: int = 0
    impulse_density: float = 3.0
    single_realization: bool = False
    isotropic_3D_sampling: bool = False
    one_d: bool = False
    scheme: int = A.SCHEME_UNI
    correlation_xy: bool = False
    correlation_context: int = A.CTX_RENEWAL_PLUS
    max_bounces: int = 1024
    # gaussian_process.covariance
    cov_kind: int = A.COV_SE
    sigma: float = 1.0
    length_scale: float = 1.0
    local_scale: float = 3.0
    aniso: tuple = (1.0, 1.0, 1.0)
    # "useAnisoMtx" and the row-major "anisoMtx" (GPFunctions.cpp:663-681)
    use_aniso_mtx: bool = False
    aniso_mtx: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    ns_multires: bool = False
    ns_ls_field: int = A.LS_CONST
    ns_ls_params: tuple = (1.0, 0.0, 0.0, 0.0)
    ns_ls_max: float = 1.0
    ns_var: float = 1.0
    # "noise" fields (ProceduralNoise): type index of A.NOISE_TYPES and the JSON min, max,
    # start, end, min2, max2, start2, end2
    ns_var_field: int = A.VAR_CONST
    ns_var_noise: int = 0
    ns_var_noise_params: tuple = (1.0, 500.0, 0.0, 1.0, 1.0, 500.0, 0.0, 1.0)
    ns_ls_noise: int = 0
    ns_ls_noise_params: tuple = (1.0, 500.0, 0.0, 1.0, 1.0, 500.0, 0.0, 1.0)
    # gaussian_process.mean
    mean_kind: int = A.MEAN_SPHERE
    mean_center: tuple = (0.0, 0.0, 0.0)
    mean_radius: float = 1.0
    mean_half: tuple = (0.5, 0.5, 0.5)
    plane_normal: tuple = (1.0, 0.0, 0.0)
    plane_offset: float = 0.0
    # TabulatedMean (GPFunctions.cpp:143-153): grid, offset, scale, volume
    mean_grid: Optional[DenseMeanGrid] = None
    mean_offset: float = 0.0
    mean_scale: float = 1.0
    mean_volume: bool = False
    # ProceduralMean (GPFunctions.cpp:250-266): SdfFunctions function, the float inverse of
    # its transform (rows 0-2, row-major), min, scale, offset
    sdf_function: int = 0
    sdf_inverse_transform: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    sdf_min: float = -3.4028234663852886e38
    sdf_scale: float = 1.0
    sdf_offset: float = 0.0
    extra: Dict = field(default_factory=dict)

    @classmethod
    def from_json(cls, js: dict, mean_grid: Optional[DenseMeanGrid] = None,
                  base_dir: Optional[str] = None) -> "SparseConvolutionNoiseMedium":
        """``mean_grid`` supplies a tabulated mean's voxels when the JSON names a grid the
        host application reads itself (a "vdb" grid); "dense" grids are read from their
        .npy file, relative to ``base_dir``."""
        if js.get("type", "sparse_conv_noise") != "sparse_conv_noise":
            raise MediumConfigError(f"not a sparse_conv_noise medium: {js.get('type')!r}")
        m = cls()
        m.step_size = float(js.get("step_size", m.step_size))
        m.min_step = int(js.get("min_step", m.min_step))
        m.seed = int(js.get("seed", m.seed)) & 0xFFFFFFFF
        m.impulse_density = float(js.get("impulse_density", m.impulse_density))
        m.single_realization = bool(js.get("single_realization", m.single_realization))
        m.isotropic_3D_sampling = bool(js.get("isotropic_3D_sampling", False))
        m.one_d = bool(js.get("1D_sampling", False))
        scheme = js.get("1D_sampling_scheme", "uni")
        if scheme not in _SCHEME:
            raise MediumConfigError(f"Invalid sparse conv sampling scheme: '{scheme}'")
        m.scheme = _SCHEME[scheme]
        m.correlation_xy = bool(js.get("1D_gradient_correlationXY", False))
        # the reference parses a placeholder and FAILs when the key is absent (GaussianProcessMedium.cpp:105-107)
        ctx = js.get("correlation_context", "goldfish")
        if ctx not in _CTX:
            raise MediumConfigError(f"Invalid correlation context: '{ctx}'")
        m.correlation_context = _CTX[ctx]
        if js.get("intersect_method", "gp_discrete") != "gp_discrete":
            raise MediumConfigError("only intersect_method 'gp_discrete' is a sparse-convolution flight")
        m.max_bounces = int(js.get("max_bounces", m.max_bounces))
        gp = js.get("gaussian_process", {})
        mean = gp.get("mean", {"type": "spherical"})
        mt = mean.get("type")
        if mt == "spherical":
            m.mean_kind = A.MEAN_SPHERE
            m.mean_center = tuple(float(x) for x in mean.get("center", (0.0, 0.0, 0.0)))
            m.mean_radius = float(mean.get("radius", 1.0))
        elif mt == "box":
            m.mean_kind = A.MEAN_BOX
            m.mean_center = tuple(float(x) for x in mean.get("center", (0.0, 0.0, 0.0)))
            h = mean.get("half_extent", 0.5)
            m.mean_half = tuple(float(x) for x in (h if isinstance(h, (list, tuple)) else (h, h, h)))
        elif mt == "plane":
            m.mean_kind = A.MEAN_PLANE
            m.plane_normal = tuple(float(x) for x in mean.get("normal", (1.0, 0.0, 0.0)))
            m.plane_offset = float(mean.get("offset", 0.0))
        elif mt == "tabulated":
            m.mean_kind = A.MEAN_GRID
            m.mean_offset = float(mean.get("offset", 0.0))
            m.mean_scale = float(mean.get("scale", 1.0))
            m.mean_volume = bool(mean.get("volume", False))
            g = mean.get("grid", {})
            if mean_grid is not None:
                m.mean_grid = mean_grid
            elif g.get("type") == "dense":
                m.mean_grid = DenseMeanGrid.from_json(g, base_dir)
            else:
                raise MediumConfigError(
                    f"tabulated mean grid of type {g.get('type')!r}: pass its dense export "
                    "(origin, dims, voxels, world_to_index, clamp box) as mean_grid")
        elif mt == "procedural":
            m.mean_kind = A.MEAN_PROCEDURAL
            if "func" in mean:
                fn = mean["func"]
            elif "f" in mean:
                raise MediumConfigError("a procedural mean over a procedural scalar \"f\" is not served; "
                                        "name an SDF function in \"func\"")
            else:
                fn = "knob"
            if fn not in A.SDF_FUNCTIONS:
                raise MediumConfigError(f"Invalid sdf function: '{fn}'")
            m.sdf_function = A.SDF_FUNCTIONS.index(fn)
            try:
                inv = mat4f.invert(mat4f.from_json(mean.get("transform")))
            except ValueError as e:
                raise MediumConfigError(f"procedural mean transform: {e}") from None
            m.sdf_inverse_transform = tuple(float(v) for v in inv[:12])
            m.sdf_min = float(np.float32(mean.get("min", m.sdf_min)))
            m.sdf_scale = float(mean.get("scale", 1.0))
            m.sdf_offset = float(mean.get("offset", 0.0))
        else:
            raise MediumConfigError(f"mean type {mt!r} is not served by the GPU flight path")
        cov = gp.get("covariance", {"type": "squared_exponential"})
        ct = cov.get("type")
        if ct == "proc_nonstationary":
            m.cov_kind = A.COV_PROC_NS
            m.ns_multires = bool(cov.get("multiResolutionGrid", False))
            ls = cov.get("ls", {})
            if ls.get("type") == "sincos":
                m.ns_ls_field = A.LS_SINCOS
                m.ns_ls_params = (float(ls["base"]), float(ls["amplitude"]), float(ls["frequency"]), 0.0)
                m.ns_ls_max = float(ls["base"]) + float(ls["amplitude"])
            elif ls.get("type", "constant") == "constant":
                v = float(ls.get("value", 1.0))
                m.ns_ls_field, m.ns_ls_params, m.ns_ls_max = A.LS_CONST, (v, 0.0, 0.0, 0.0), v
            elif ls.get("type") == "noise":
                m.ns_ls_field = A.LS_NOISE
                m.ns_ls_noise, m.ns_ls_noise_params = _noise_field(ls)
                m.ns_ls_max = _noise_max(m.ns_ls_noise, m.ns_ls_noise_params)
            else:
                raise MediumConfigError(f"length-scale field {ls.get('type')!r} is not served by the GPU flight path")
            var = cov.get("var", 1.0)
            if isinstance(var, dict):
                if var.get("type") == "noise":
                    m.ns_var_field = A.VAR_NOISE
                    m.ns_var_noise, m.ns_var_noise_params = _noise_field(var)
                    var = 1.0
                elif var.get("type", "constant") != "constant":
                    raise MediumConfigError(f"variance field {var.get('type')!r} is not served by the GPU flight path")
                else:
                    var = var.get("value", 1.0)
            m.ns_var = float(var)
            cov = cov.get("cov", {"type": "squared_exponential"})
            ct = cov.get("type")
            if cov.get("useAnisoMtx", False):
                raise MediumConfigError("useAnisoMtx is served for the stationary squared_exponential "
                                        "covariance, not inside proc_nonstationary")
        if ct != "squared_exponential":
            raise MediumConfigError(f"covariance {ct!r} is not served by the GPU flight path")
        m.sigma = float(cov.get("sigma", 1.0))
        m.length_scale = float(cov.get("lengthScale", 1.0))
        m.local_scale = float(cov.get("localScale", 3.0))
        a = cov.get("aniso", (1.0, 1.0, 1.0))
        m.aniso = tuple(float(x) for x in (a if isinstance(a, (list, tuple)) else (a, a, a)))
        if cov.get("useAnisoMtx", False):
            mt = cov.get("anisoMtx")
            # JsonPtr::get(Eigen::Matrix3f): a 9-element array, row-major
            if not isinstance(mt, (list, tuple)) or len(mt) != 9:
                raise MediumConfigError("useAnisoMtx needs \"anisoMtx\": an array of 9 numbers, row-major")
            m.use_aniso_mtx = True
            m.aniso_mtx = tuple(float(np.float32(x)) for x in mt)
            if m.one_d:
                raise MediumConfigError("useAnisoMtx is served for 3D sampling (world space or the "
                                        "isotropic ray space), not 1D_sampling")
        return m

    def to_desc(self) -> A.FieldDesc:
        d = A.FieldDesc()
        d.cov_kind = self.cov_kind
        d.sigma, d.length_scale, d.kernel_scale = self.sigma, self.length_scale, self.local_scale
        for i in range(3):
            d.aniso[i] = self.aniso[i]
            d.mean_center[i] = self.mean_center[i]
            d.mean_half[i] = self.mean_half[i]
            d.plane_normal[i] = self.plane_normal[i]
        d.ns_multires = int(self.ns_multires)
        d.ns_ls_field = self.ns_ls_field
        for i in range(4):
            d.ns_ls_params[i] = self.ns_ls_params[i]
        d.ns_ls_max, d.ns_var = self.ns_ls_max, self.ns_var
        d.ns_var_field, d.ns_var_noise, d.ns_ls_noise = self.ns_var_field, self.ns_var_noise, self.ns_ls_noise
        for i in range(8):
            d.ns_var_noise_params[i] = self.ns_var_noise_params[i]
            d.ns_ls_noise_params[i] = self.ns_ls_noise_params[i]
        d.mean_kind, d.mean_radius, d.plane_offset = self.mean_kind, self.mean_radius, self.plane_offset
        d.global_seed = self.seed
        d.impulse_density = self.impulse_density
        d.single_realization = int(self.single_realization)
        d.ctxt = self.correlation_context
        d.one_d = int(self.one_d)
        d.scheme = self.scheme
        d.correlation_xy = int(self.correlation_xy)
        d.step_size, d.min_step, d.max_bounce = self.step_size, self.min_step, self.max_bounces
        d.isotropic_3d = int(self.isotropic_3D_sampling and not self.one_d)
        d.sdf_function = self.sdf_function
        for i in range(12):
            d.sdf_inverse_transform[i] = self.sdf_inverse_transform[i]
        d.sdf_min, d.sdf_scale, d.sdf_offset = self.sdf_min, self.sdf_scale, self.sdf_offset
        d.use_aniso_mtx = int(self.use_aniso_mtx)
        for i in range(9):
            d.aniso_mtx[i] = self.aniso_mtx[i]
        return d

    @property
    def kernel_radius(self) -> float:
        """World cell size of the 3D grid, splattingKernelRadius(false, 1) (GPFunctions.cpp:706-719)."""
        f = np.float32
        l_conv = f(self.length_scale) * np.sqrt(f(2.0)) / f(2)
        if self.use_aniso_mtx:
            # the largest component of the row sums of l_conv * anisoMtx (:714-717)
            L = [f(l_conv) * f(a) for a in self.aniso_mtx]
            la = [(L[3 * i] + L[3 * i + 1]) + L[3 * i + 2] for i in range(3)]
        else:
            la = [f(l_conv) * f(a) for a in self.aniso]
        ls = f(1.0) if self.cov_kind == A.COV_SE or self.ns_multires else f(self.ns_ls_max)
        return float(f(self.local_scale) * ls * max(la))

    def compute_seed(self, pss: np.ndarray, scene_seed) -> np.ndarray:
        """SparseConvolutionNoiseRealization::computeSeed (SparseConvolutionNoise.cpp:57-66), vectorised."""
        pss = np.asarray(pss, np.uint32).reshape(-1, 4)
        n = pss.shape[0]
        if self.single_realization:
            return np.full(n, self.seed, np.uint32)
        w = np.zeros(n, np.uint32) if self.correlation_context == A.CTX_GLOBAL else pss[:, 3]
        with np.errstate(over="ignore"):
            return (np.uint32(self.seed) + xxhash32_4(pss[:, 0], pss[:, 1], pss[:, 2], w)
                    + xxhash32_1(np.broadcast_to(np.asarray(scene_seed, np.uint32), (n,)))).astype(np.uint32)

    @staticmethod
    def normal_seed(pss: np.ndarray, scene_seed) -> np.ndarray:
        """xxhash32(pixelSampleSegment) + xxhash32(sceneSeed) + 1 (SparseConvolutionNoiseMedium.cpp:237-238)."""
        pss = np.asarray(pss, np.uint32).reshape(-1, 4)
        n = pss.shape[0]
        with np.errstate(over="ignore"):
            return (xxhash32_4(pss[:, 0], pss[:, 1], pss[:, 2], pss[:, 3])
                    + xxhash32_1(np.broadcast_to(np.asarray(scene_seed, np.uint32), (n,))) + np.uint32(1)).astype(np.uint32)


RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "jitter", "noise_seed", "normal_seed",
              "flags", "prev_value", "prev_gx", "prev_gy", "prev_gz")


class RayBatch:
    """Structure-of-arrays wavefront of flight requests resident in device memory."""

    def __init__(self, tensors: Dict[str, "object"]):
        self.t = tensors
        self.n = int(tensors["ox"].shape[0])

    @classmethod
    def from_numpy(cls, arrays: Dict[str, np.ndarray], device="cuda"):
        """`arrays` holds RAY_FIELDS and, for 1D media with a per-ray SE length
        scale, an optional float32 "length_scale"."""
        import torch
        out = {}
        for k in RAY_FIELDS + ("length_scale",):
            if k not in arrays:
                continue
            a = np.ascontiguousarray(arrays[k])
            if a.dtype == np.uint32:
                a = a.view(np.int32)
            out[k] = torch.from_numpy(a).to(device)
        return cls(out)

    def soa(self) -> A.RaySoA:
        s = A.RaySoA()
        for k in RAY_FIELDS:
            setattr(s, k, self.t[k].data_ptr())
        ls = self.t.get("length_scale")
        s.length_scale = None if ls is None else ls.data_ptr()
        return s


class GpisDevice:
    """One gpis_ctx on the current HIP device."""

    def __init__(self, medium: SparseConvolutionNoiseMedium):
        import torch
        if not torch.cuda.is_available():
            raise A.GpisError("no HIP device visible")
        self.medium = medium
        self.L = A.lib()
        self.desc = medium.to_desc()
        h = C.c_void_p()
        A.check(self.L.gpis_create(C.byref(self.desc), C.byref(h)), None, "gpis_create")
        self.h = h
        if medium.mean_kind == A.MEAN_GRID:
            if medium.mean_grid is None:
                raise A.GpisError("tabulated mean without a grid")
            g = medium.mean_grid.to_c(medium.mean_offset, medium.mean_scale, medium.mean_volume)
            A.check(self.L.gpis_set_mean_grid(self.h, C.byref(g)), self.h, "gpis_set_mean_grid")

    def close(self):
        if getattr(self, "h", None):
            self.L.gpis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def set_domain(self, lo, hi, seed: Optional[int] = None):
        lo3 = (C.c_float * 3)(*lo)
        hi3 = (C.c_float * 3)(*hi)
        s = self.medium.seed if seed is None else seed
        A.check(self.L.gpis_set_domain(self.h, lo3, hi3, s & 0xFFFFFFFF), self.h, "gpis_set_domain")

    def set_launch(self, waves_per_cu=0, block_size=0):
        A.check(self.L.gpis_set_launch(self.h, waves_per_cu, block_size), self.h, "gpis_set_launch")

    def set_refill(self, batch: int):
        """Lanes per wave served together by the finish/bounce/refill step (1..64)."""
        A.check(self.L.gpis_set_refill(self.h, batch), self.h, "gpis_set_refill")

    def set_coop_cells(self, on: bool):
        """Wave-cooperative generated cells in the 3D flight kernels (on by default where the
        medium generates its cells per flight); results are the same either way."""
        A.check(self.L.gpis_set_coop_cells(self.h, 1 if on else 0), self.h, "gpis_set_coop_cells")

    def set_lists(self, sub: int, points: int = 2):
        """Support lists of the impulse table: sub-cells per cell edge (1..8, 0 for none) and
        evaluation points per list pass (1 or 2)."""
        A.check(self.L.gpis_set_lists(self.h, sub, points), self.h, "gpis_set_lists")

    def set_precision(self, mode: str):
        """Arithmetic of the 1D flights: "fast" (FP32 splats and analytic mean, the default)
        or "exact" (the reference's rounding step by step, bit-identical with the oracle)."""
        m = {"fast": 0, "exact": 1}.get(mode)
        if m is None:
            raise MediumConfigError(f"unknown precision {mode!r} (fast, exact)")
        A.check(self.L.gpis_set_precision(self.h, m), self.h, "gpis_set_precision")

    def impulse_table(self, lo, dims, seed):
        import torch
        n = int(self.medium.impulse_density)
        cells = int(dims[0]) * int(dims[1]) * int(dims[2])
        out = torch.empty(cells * n * 4, dtype=torch.float32, device="cuda")
        lo3 = (C.c_int32 * 3)(*[int(x) for x in lo])
        d3 = (C.c_int32 * 3)(*[int(x) for x in dims])
        A.check(self.L.gpis_impulse_table(self.h, lo3, d3, seed & 0xFFFFFFFF, out.data_ptr(), self._stream()),
                self.h, "gpis_impulse_table")
        return out.view(cells, n, 4)

    def eval_points(self, x, y, z, seed):
        import torch
        n = int(x.shape[0])
        v = torch.empty(n, dtype=torch.float32, device=x.device)
        gx, gy, gz = torch.empty_like(v), torch.empty_like(v), torch.empty_like(v)
        A.check(self.L.gpis_eval_points(self.h, x.data_ptr(), y.data_ptr(), z.data_ptr(), n, seed & 0xFFFFFFFF,
                                        v.data_ptr(), gx.data_ptr(), gy.data_ptr(), gz.data_ptr(), self._stream()),
                self.h, "gpis_eval_points")
        return v, gx, gy, gz

    def alloc_reply(self, n, device="cuda"):
        import torch
        f = dict(dtype=torch.float32, device=device)
        return {"t": torch.empty(n, **f), "value": torch.empty(n, **f), "gx": torch.empty(n, **f),
                "gy": torch.empty(n, **f), "gz": torch.empty(n, **f),
                "outcome": torch.empty(n, dtype=torch.int32, device=device),
                "evals": torch.empty(n, dtype=torch.int32, device=device)}

    def trace_flights(self, rays: RayBatch, reply=None):
        if reply is None:
            reply = self.alloc_reply(rays.n)
        r = A.ReplySoA()
        for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
            setattr(r, k, reply[k].data_ptr())
        soa = rays.soa()
        A.check(self.L.gpis_trace_flights(self.h, C.byref(soa), rays.n, C.byref(r), self._stream()),
                self.h, "gpis_trace_flights")
        return reply

    def trace_flights_1d(self, rays: RayBatch, light=None, reply=None):
        """One 1D flight per ray (1D_sampling media).  With `light` = (x, y, z), each
        hit also returns the NEE gradient and pdf toward that point light in
        reply["nee_gx"/"nee_gy"/"nee_gz"/"nee_pdf"]."""
        import torch
        if reply is None:
            reply = self.alloc_reply(rays.n)
        r = A.ReplySoA()
        for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
            setattr(r, k, reply[k].data_ptr())
        soa = rays.soa()
        lp, nee = None, None
        if light is not None:
            for k in ("nee_gx", "nee_gy", "nee_gz", "nee_pdf"):
                if k not in reply:
                    reply[k] = torch.empty(rays.n, dtype=torch.float32, device=reply["t"].device)
            nee = A.NeeOut(reply["nee_gx"].data_ptr(), reply["nee_gy"].data_ptr(), reply["nee_gz"].data_ptr(),
                           reply["nee_pdf"].data_ptr())
            lp = (C.c_float * 3)(*[float(x) for x in light])
        A.check(self.L.gpis_trace_flights_1d(self.h, C.byref(soa), rays.n, C.byref(r), lp,
                                             C.byref(nee) if nee is not None else None, self._stream()),
                self.h, "gpis_trace_flights_1d")
        return reply

    @staticmethod
    def path_seeds(pss, device="cuda"):
        """Upload each path's pixelSampleSegment (n x 4, uint32) as the four device
        columns gpis_trace_paths_1d reads."""
        import torch
        p = np.ascontiguousarray(np.asarray(pss, np.uint32).reshape(-1, 4))
        return [torch.from_numpy(np.ascontiguousarray(p[:, i]).view(np.int32)).to(device) for i in range(4)]

    def alloc_path1d_reply(self, n, device="cuda"):
        import torch
        f = dict(dtype=torch.float32, device=device)
        i = dict(dtype=torch.int32, device=device)
        reply = {k: torch.empty(n, **f) for k in ("t0", "gx0", "gy0", "gz0", "nee", "t_last")}
        reply.update({k: torch.empty(n, **i) for k in ("outcome0", "flights", "hits", "shadows", "visible",
                                                       "outcome_last", "evals")})
        return reply

    def trace_paths_1d(self, rays: RayBatch, pss, scene_seed: int, max_bounces: int, phase: str = "mirror",
                       bounce_seed: int = 7, box: float = 1.0, light=None, reply=None):
        """Device-resident paths of a 1D_sampling medium (gpis_trace_paths_1d): each ray
        starts a path whose hits run the hit-time phase ("mirror" or "lambert" BSDF of
        albedo 1) and, with `light`, the light sample with its shadow flight toward that
        point light.  `pss` holds each path's pixelSampleSegment, as an n x 4 uint32
        array or as the device columns of path_seeds(): flight b uses (px, py, spp,
        w + b) and `scene_seed` for its seeds."""
        kind = {"mirror": A.PHASE_MIRROR, "lambert": A.PHASE_LAMBERT}.get(phase)
        if kind is None:
            raise MediumConfigError(f"unknown phase {phase!r} (mirror, lambert)")
        dev = rays.t["ox"].device
        cols = pss if isinstance(pss, (list, tuple)) else self.path_seeds(pss, dev)
        if len(cols) != 4 or any(int(c.shape[0]) != rays.n for c in cols):
            raise A.GpisError("pss needs four device columns of one entry per path")
        seeds = A.PathSeeds(cols[0].data_ptr(), cols[1].data_ptr(), cols[2].data_ptr(), cols[3].data_ptr(),
                            int(scene_seed) & 0xFFFFFFFF)
        if reply is None:
            reply = self.alloc_path1d_reply(rays.n, dev)
        r = A.Path1DReply()
        for k, _ in A.Path1DReply._fields_:
            setattr(r, k, reply[k].data_ptr())
        soa = rays.soa()
        lp = None if light is None else (C.c_float * 3)(*[float(x) for x in light])
        A.check(self.L.gpis_trace_paths_1d(self.h, C.byref(soa), C.byref(seeds), rays.n, max_bounces, kind,
                                           bounce_seed & 0xFFFFFFFF, box, lp, C.byref(r), self._stream()),
                self.h, "gpis_trace_paths_1d")
        reply["_seeds"] = cols  # device inputs stay referenced until the caller synchronises
        return reply

    def alloc_path_reply(self, n, device="cuda"):
        import torch
        f = dict(dtype=torch.float32, device=device)
        i = dict(dtype=torch.int32, device=device)
        return {"t0": torch.empty(n, **f), "gx0": torch.empty(n, **f), "gy0": torch.empty(n, **f),
                "gz0": torch.empty(n, **f), "outcome0": torch.empty(n, **i), "flights": torch.empty(n, **i),
                "hits": torch.empty(n, **i), "evals": torch.empty(n, **i), "t_last": torch.empty(n, **f),
                "outcome_last": torch.empty(n, **i)}

    def trace_paths(self, rays: RayBatch, max_bounces: int, bounce_seed: int, box: float, reply=None, pss=None,
                    scene_seed: int = 0):
        """Paths of up to 1 + max_bounces 3D flights with diffuse bounces (gpis_trace_paths).
        `pss` (optional): each path's pixelSampleSegment, as an n x 4 uint32 array or the
        device columns of path_seeds(); flight b then uses the realization seed of (px, py,
        spp, w + b) and `scene_seed`.  Without it every flight uses the ray's noise_seed."""
        if reply is None:
            reply = self.alloc_path_reply(rays.n)
        r = A.PathReply()
        for k, _ in A.PathReply._fields_:
            setattr(r, k, reply[k].data_ptr() if k in reply else None)
        soa = rays.soa()
        seeds = None
        if pss is not None:
            dev = rays.t["ox"].device
            cols = pss if isinstance(pss, (list, tuple)) else self.path_seeds(pss, dev)
            if len(cols) != 4 or any(int(c.shape[0]) != rays.n for c in cols):
                raise A.GpisError("pss needs four device columns of one entry per path")
            seeds = A.PathSeeds(cols[0].data_ptr(), cols[1].data_ptr(), cols[2].data_ptr(), cols[3].data_ptr(),
                                int(scene_seed) & 0xFFFFFFFF)
            reply["_seeds"] = cols  # device inputs stay referenced until the caller synchronises
        A.check(self.L.gpis_trace_paths(self.h, C.byref(soa), C.byref(seeds) if seeds is not None else None,
                                        rays.n, max_bounces, bounce_seed & 0xFFFFFFFF, box, C.byref(r),
                                        self._stream()), self.h, "gpis_trace_paths")
        return reply

    def counters(self) -> Dict[str, int]:
        c = A.Counters()
        A.check(self.L.gpis_get_counters(self.h, C.byref(c)), self.h, "gpis_get_counters")
        return {k: int(getattr(c, k)) for k, _ in A.Counters._fields_}

    def reset_counters(self):
        A.check(self.L.gpis_reset_counters(self.h), self.h, "gpis_reset_counters")

    # ---- Medium-shaped views of a reply ---------------------------------
    @staticmethod
    def sample_distance(reply) -> Dict[str, "object"]:
        """sampleDistance semantics per flight: ok (not Failed), exited (no hit), t, aniso."""
        oc = reply["outcome"] & 0xFF
        return {"ok": (oc == A.HIT) | (oc == A.EXIT), "exited": oc == A.EXIT, "t": reply["t"],
                "aniso": (reply["gx"], reply["gy"], reply["gz"])}

    @staticmethod
    def transmittance(reply):
        """transmittance semantics: 1 when the flight exits, 0 at any crossing (GaussianProcessMedium.cpp:394-421)."""
        oc = reply["outcome"] & 0xFF
        return (oc == A.EXIT).float()
