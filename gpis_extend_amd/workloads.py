"""Synthetic workloads of the sparse-convolution GPIS flight benchmarks.

Each builder returns the medium (reference JSON form) and host numpy arrays of
the flight requests.  Random streams are numpy PCG64 generators with the
committed seeds named below, so every run sees the same inputs.
"""
from __future__ import annotations

import math

import numpy as np

from . import _capi as A
from .medium import DenseMeanGrid, SparseConvolutionNoiseMedium

SEED_POINTS = 0x5EED0000
SEED_RAYS = 0x5EED0001
SEED_CAMERA = 0x5EED0002
SEED_BOX = 0x5EED0003
SEED_RENEWAL1D = 0x5EED0004
LIGHT_POS = (0.0, 2.0, 0.0)


def config0_json(length_scale=0.1):
    """Stationary SE field: sphere SDF mean (r 0.5), sigma 1, fixed 8 Bernoulli impulses
    per cell, single realization with global seed 1."""
    return {
        "type": "sparse_conv_noise", "max_bounces": 1024,
        "gaussian_process": {
            "type": "standard",
            "mean": {"type": "spherical", "center": [0.0, 0.0, 0.0], "radius": 0.5},
            "covariance": {"type": "squared_exponential", "sigma": 1.0, "lengthScale": length_scale},
        },
        "step_size": 0.01, "min_step": 8, "seed": 1, "impulse_density": 8,
        "single_realization": True, "correlation_context": "global", "intersect_method": "gp_discrete",
        "isotropic_3D_sampling": False, "1D_sampling": False,
    }


def config1_json():
    return config0_json(length_scale=0.05)


def config2_json():
    """Nonstationary multi-resolution field, box SDF mean (half extent 0.6), Renewal+."""
    return {
        "type": "sparse_conv_noise", "max_bounces": 1024,
        "gaussian_process": {
            "type": "standard",
            "mean": {"type": "box", "center": [0.0, 0.0, 0.0], "half_extent": 0.6},
            "covariance": {
                "type": "proc_nonstationary", "multiResolutionGrid": True,
                "ls": {"type": "sincos", "base": 0.02, "amplitude": 0.18, "frequency": 4.0 * math.pi},
                "cov": {"type": "squared_exponential", "sigma": 1.0, "lengthScale": 1.0},
            },
        },
        "step_size": 0.01, "min_step": 8, "seed": 0, "impulse_density": 8,
        "single_realization": False, "correlation_context": "renewal+", "intersect_method": "gp_discrete",
        "isotropic_3D_sampling": False, "1D_sampling": False,
    }


def config3_json():
    """1D along-ray SE process (sigma 1) over the sphere SDF mean of the stationary
    field (r 0.5), Renewal+ without xy correlation, MIS scheme.  Each ray carries its
    own SE length scale; the medium's lengthScale is the default for rays without one."""
    return {
        "type": "sparse_conv_noise", "max_bounces": 1024,
        "gaussian_process": {
            "type": "standard",
            "mean": {"type": "spherical", "center": [0.0, 0.0, 0.0], "radius": 0.5},
            "covariance": {"type": "squared_exponential", "sigma": 1.0, "lengthScale": 0.1},
        },
        "step_size": 0.01, "min_step": 8, "seed": 3, "impulse_density": 8,
        "single_realization": False, "correlation_context": "renewal+", "intersect_method": "gp_discrete",
        "isotropic_3D_sampling": False, "1D_sampling": True, "1D_sampling_scheme": "mis",
        "1D_gradient_correlationXY": False,
    }


def with_tabulated_mean(js, offset=0.0, scale=1.0, volume=False):
    """The medium `js` with its mean replaced by a TabulatedMean over a VDB grid, in the
    reference's scene form (data/example-scenes/lattice-cubic-gpis/shaderball.json); the
    grid's voxels come separately, as a DenseMeanGrid."""
    import copy
    out = copy.deepcopy(js)
    out["gaussian_process"]["mean"] = {
        "type": "tabulated", "volume": bool(volume), "scale": float(scale), "offset": float(offset),
        "grid": {"type": "vdb", "file": "sdf.vdb", "density_name": "sdf", "interpolate": "linear"}}
    return out


def with_procedural_mean(js, func="knob", transform=None, scale=1.0, offset=0.0, minimum=None):
    """The medium `js` with its mean replaced by a ProceduralMean over an SdfFunctions
    function, in the reference's scene form (data/example-scenes/two-spheres-backlight/
    scene.json: func "two_spheres", transform {"scale": 0.1}, scale 0.1)."""
    import copy
    out = copy.deepcopy(js)
    mean = {"type": "procedural", "func": func, "scale": float(scale), "offset": float(offset)}
    if transform is not None:
        mean["transform"] = transform
    if minimum is not None:
        mean["min"] = float(minimum)
    out["gaussian_process"]["mean"] = mean
    return out


def sphere_sdf_grid(radius=0.5, center=(0.0, 0.0, 0.0), half=1.06, spacing=0.02):
    """Signed distance to a sphere tabulated on [-half, half]^3 with the given voxel spacing
    (the dense export of an SDF grid like the reference's shaderball_sdf.vdb)."""
    c = np.asarray(center, np.float64)

    def sdf(x, y, z):
        return np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - radius
    return DenseMeanGrid.sample(sdf, (-half,) * 3, (half,) * 3, spacing)


def points_workload(n=1 << 20, seed=SEED_POINTS):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)


def _jitter(rng, n):
    """Uniform floats in [0, 1) formed as UniformSampler::next1D forms them from a
    32-bit draw (BitManip::normalizedUint): never 1.0 after rounding."""
    x = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    return (((x >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)).astype(np.float32)


def _empty_rays(n):
    return {
        "ox": np.zeros(n, np.float32), "oy": np.zeros(n, np.float32), "oz": np.zeros(n, np.float32),
        "dx": np.zeros(n, np.float32), "dy": np.zeros(n, np.float32), "dz": np.zeros(n, np.float32),
        "near_t": np.zeros(n, np.float32), "far_t": np.zeros(n, np.float32), "jitter": np.zeros(n, np.float32),
        "noise_seed": np.zeros(n, np.uint32), "normal_seed": np.zeros(n, np.uint32),
        "flags": np.ones(n, np.uint32), "prev_value": np.zeros(n, np.float32),
        "prev_gx": np.zeros(n, np.float32), "prev_gy": np.zeros(n, np.float32), "prev_gz": np.zeros(n, np.float32),
    }


def _fill_seeds(rays, medium: SparseConvolutionNoiseMedium, pss, scene_seed):
    rays["noise_seed"] = medium.compute_seed(pss, scene_seed)
    rays["normal_seed"] = medium.normal_seed(pss, scene_seed)
    rays["pss"] = pss
    rays["scene_seed"] = scene_seed


def sphere_rays(medium, n=65536, seed=SEED_RAYS, radius=2.0, scene_seed=0):
    """Origins uniform on |x| = radius, directions toward uniform points of the unit
    ball; the flight runs to the far side of the radius sphere (near 0, far = chord)."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    u = rng.normal(size=(n, 3))
    u = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(0, 1, size=(n, 1)) ** (1.0 / 3.0)
    d = u - o
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    # re-normalise in float so |d| is within an ulp of 1
    d = d / np.sqrt((d.astype(np.float64) ** 2).sum(1, keepdims=True)).astype(np.float32)
    rays = _empty_rays(n)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    rays["far_t"] = (-2.0 * (o.astype(np.float64) * d).sum(1)).astype(np.float32)
    rays["jitter"] = _jitter(rng, n)
    pss = np.zeros((n, 4), np.uint32)
    pss[:, 0] = np.arange(n, dtype=np.uint32)
    _fill_seeds(rays, medium, pss, scene_seed)
    return rays


def box_interval(o, d, b=1.0):
    """Entry/exit distances of rays o + t d through the box [-b, b]^3 (slab method)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t1 = (-b - o) * inv
        t2 = (b - o) * inv
    tn = np.nanmax(np.minimum(t1, t2), axis=1)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    return np.maximum(tn, 0.0), tf


def camera_rays(medium, width=2048, height=2048, fov_deg=40.0, eye=(0.0, 0.0, 3.0), seed=SEED_CAMERA,
                scene_seed=0, box=1.0):
    """Pinhole camera at `eye` looking at the origin; one ray per pixel centre.  Rays
    that miss the box [-box, box]^3 are dropped (they issue no flight).  Rays are
    ordered in 8x8 pixel tiles so a wavefront holds neighbouring pixels."""
    rng = np.random.default_rng(seed)
    th = math.tan(math.radians(fov_deg) * 0.5)
    aspect = width / height
    # tile order: 8x8 tiles, row-major inside a tile
    ty, tx = np.meshgrid(np.arange(height // 8), np.arange(width // 8), indexing="ij")
    iy, ix = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    px = (tx.reshape(-1, 1) * 8 + ix.reshape(1, -1)).reshape(-1)
    py = (ty.reshape(-1, 1) * 8 + iy.reshape(1, -1)).reshape(-1)
    sx = ((px + 0.5) / width * 2.0 - 1.0) * th * aspect
    sy = (1.0 - (py + 0.5) / height * 2.0) * th
    e = np.asarray(eye, np.float64)
    fwd = -e / np.linalg.norm(e)
    up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    up2 = np.cross(right, fwd)
    d = fwd[None, :] + sx[:, None] * right[None, :] + sy[:, None] * up2[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(e, d.shape)
    tn, tf = box_interval(o, d, box)
    keep = tf > tn
    d = d[keep].astype(np.float32)
    d = d / np.sqrt((d.astype(np.float64) ** 2).sum(1, keepdims=True)).astype(np.float32)
    n = d.shape[0]
    rays = _empty_rays(n)
    rays["ox"][:] = e[0]; rays["oy"][:] = e[1]; rays["oz"][:] = e[2]
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    rays["near_t"] = tn[keep].astype(np.float32)
    rays["far_t"] = tf[keep].astype(np.float32)
    rays["jitter"] = _jitter(rng, n)
    pss = np.zeros((n, 4), np.uint32)
    pss[:, 0] = px[keep]
    pss[:, 1] = py[keep]
    _fill_seeds(rays, medium, pss, scene_seed)
    rays["pixel"] = np.stack([px[keep], py[keep]], 1).astype(np.int32)
    return rays


def box_surface_rays(medium, n=1 << 20, half=0.6, seed=SEED_BOX, scene_seed=2, box=1.0):
    """Origins uniform on the surface of the box SDF (half extent `half`), directions
    cosine-weighted about the face normal.  The first flight is conditioned on value 0
    and gradient = the face normal; per-flight seeds use Vec4u(rayId, 0, 0, flight)."""
    rng = np.random.default_rng(seed)
    face = rng.integers(0, 6, size=n)
    axis = face // 2
    sgn = np.where(face % 2 == 0, 1.0, -1.0)
    p = rng.uniform(-half, half, size=(n, 3))
    p[np.arange(n), axis] = sgn * half
    nrm = np.zeros((n, 3))
    nrm[np.arange(n), axis] = sgn
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    r, ph = np.sqrt(u1), 2 * np.pi * u2
    local = np.stack([r * np.cos(ph), r * np.sin(ph), np.sqrt(1 - u1)], 1)
    # orthonormal basis about the normal
    a = np.where(np.abs(nrm[:, :1]) > 0.5, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(nrm, a); t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    d = local[:, :1] * t + local[:, 1:2] * b + local[:, 2:] * nrm
    d = d.astype(np.float32)
    d = d / np.sqrt((d.astype(np.float64) ** 2).sum(1, keepdims=True)).astype(np.float32)
    o = p.astype(np.float32)
    rays = _empty_rays(n)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    _, tf = box_interval(o.astype(np.float64), d.astype(np.float64), box)
    rays["far_t"] = tf.astype(np.float32)
    rays["jitter"] = _jitter(rng, n)
    rays["flags"] = np.zeros(n, np.uint32)
    rays["prev_value"] = np.zeros(n, np.float32)
    rays["prev_gx"] = nrm[:, 0].astype(np.float32)
    rays["prev_gy"] = nrm[:, 1].astype(np.float32)
    rays["prev_gz"] = nrm[:, 2].astype(np.float32)
    pss = np.zeros((n, 4), np.uint32)
    pss[:, 0] = np.arange(n, dtype=np.uint32)
    _fill_seeds(rays, medium, pss, scene_seed)
    return rays


def renewal1d_rays(medium, n=1 << 21, seed=SEED_RENEWAL1D, scene_seed=3, box=1.0, ls_range=(0.02, 0.2),
                   first_id=0):
    """Secondary 1D flights: origins uniform in the shell 0.55 <= |x| <= 0.95 around the
    sphere mean, directions toward uniform points of the ball of radius 0.5 (the
    flight runs to the exit of [-box, box]^3).  Each ray is conditioned on a value
    v0 ~ N(0.3, 0.2^2) and a gradient whose derivative along the ray is ~ N(0, 1)
    (prev_grad = g_d * dir), with its own SE length scale l ~ U(ls_range).  `seed` is an
    integer or a numpy SeedSequence; ray ids (the realization seeds) start at first_id."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    rad = (0.55 ** 3 + rng.uniform(size=(n, 1)) * (0.95 ** 3 - 0.55 ** 3)) ** (1.0 / 3.0)
    o *= rad
    u = rng.normal(size=(n, 3))
    u = u / np.linalg.norm(u, axis=1, keepdims=True) * 0.5 * rng.uniform(0, 1, size=(n, 1)) ** (1.0 / 3.0)
    d = u - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    d = d / np.sqrt((d.astype(np.float64) ** 2).sum(1, keepdims=True)).astype(np.float32)
    rays = _empty_rays(n)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    _, tf = box_interval(o.astype(np.float64), d.astype(np.float64), box)
    rays["far_t"] = tf.astype(np.float32)
    rays["jitter"] = _jitter(rng, n)
    rays["flags"] = np.zeros(n, np.uint32)
    rays["prev_value"] = rng.normal(0.3, 0.2, size=n).astype(np.float32)
    gd = rng.normal(0.0, 1.0, size=n).astype(np.float32)
    rays["prev_gx"] = (gd * d[:, 0]).astype(np.float32)
    rays["prev_gy"] = (gd * d[:, 1]).astype(np.float32)
    rays["prev_gz"] = (gd * d[:, 2]).astype(np.float32)
    rays["length_scale"] = rng.uniform(ls_range[0], ls_range[1], size=n).astype(np.float32)
    pss = np.zeros((n, 4), np.uint32)
    pss[:, 0] = (first_id + np.arange(n, dtype=np.int64)).astype(np.uint32)
    _fill_seeds(rays, medium, pss, scene_seed)
    return rays


def renewal1d_image_rays(medium, width=512, spp=16, seed=12, scene_seed=3):
    """The renewal workload arranged as an image: width^2 pixels, each one conditioned
    flight configuration of the renewal workload (origin, direction, bound, value,
    derivative, length scale), traced spp times with independent realizations
    (pss = (pixel, sample, 0, 0)) and jitters.  Ray index = pixel * spp + sample."""
    ss = np.random.SeedSequence(seed)
    s_cfg, s_jit = ss.spawn(2)
    npx = width * width
    base = renewal1d_rays(medium, n=npx, seed=s_cfg, scene_seed=scene_seed)
    rays = {k: np.repeat(v, spp, axis=0) for k, v in base.items() if isinstance(v, np.ndarray) and k != "pss"}
    n = npx * spp
    rays["jitter"] = _jitter(np.random.default_rng(s_jit), n)
    pss = np.zeros((n, 4), np.uint32)
    pss[:, 0] = np.repeat(np.arange(npx, dtype=np.uint32), spp)
    pss[:, 1] = np.tile(np.arange(spp, dtype=np.uint32), npx)
    _fill_seeds(rays, medium, pss, scene_seed)
    return rays


def rays_subset(rays, idx):
    return {k: (v[idx] if isinstance(v, np.ndarray) else v) for k, v in rays.items()}
