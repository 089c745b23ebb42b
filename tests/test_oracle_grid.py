"""The oracle's tabulated mean (TabulatedMean over a dense voxel grid) pinned by what
the mathematics fixes: voxel values at lattice points, exact reproduction of
trilinear functions (a transposed axis, a dropped origin or a wrong transform row
fails it), the clamp box, an independent z-first double-precision trilinear
interpolation (OpenVDB's BoxSampler order, which the reference's own GPU test holds
its dense grid to within 2e-4, src/gpu/tests/VdbMeanGridTest.cpp:84), the
offset/scale/volume rules and the forward-difference gradient
(GPFunctions.cpp:165-190), and whole flights over a finely tabulated sphere SDF
against the analytic sphere mean."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

NTH = min(8, os.cpu_count() or 1)


def _desc(vox, origin=(0, 0, 0), w2i=(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0), cmin=None, cmax=None, **kw):
    vox = np.asarray(vox, np.float32)
    nz, ny, nx = vox.shape
    cmin = (float(origin[0]), float(origin[1]), float(origin[2])) if cmin is None else cmin
    cmax = (float(origin[0] + nx - 1), float(origin[1] + ny - 1), float(origin[2] + nz - 1)) if cmax is None else cmax
    d = po.make_desc(mean_kind=po.MEAN_GRID)
    po.set_grid(d, vox, origin, w2i, cmin, cmax, kw.get("offset", 0.0), kw.get("scale", 1.0), kw.get("volume", False))
    return d


def _trilinear_fn(c):
    def f(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z
                + c[7] * x * y * z)
    return f


def test_lattice_points_return_the_voxel():
    rng = np.random.default_rng(1)
    vox = rng.normal(size=(5, 6, 7)).astype(np.float32)
    d = _desc(vox, origin=(-3, 2, 10))
    idx = np.stack(np.meshgrid(np.arange(7), np.arange(6), np.arange(5), indexing="ij"), -1).reshape(-1, 3)
    p = (idx + np.array([-3, 2, 10])).astype(np.float32)
    got = po.grid_density(d, p)
    want = vox[idx[:, 2], idx[:, 1], idx[:, 0]]
    assert np.array_equal(got, want)


def test_trilinear_functions_are_reproduced():
    """A trilinear function sampled on the lattice is interpolated exactly (up to float
    rounding) anywhere inside; the origin offset and a scaled, translated world->index
    map are part of the check."""
    rng = np.random.default_rng(2)
    c = rng.normal(size=8)
    spacing, lo = 0.25, np.array([-1.0, 0.5, -2.0])
    nx, ny, nz = 9, 7, 8
    org = np.array([4, -2, 3])
    # world point of index (i, j, k): lo + spacing * (ijk - origin)
    ii = np.arange(nx) + org[0]
    jj = np.arange(ny) + org[1]
    kk = np.arange(nz) + org[2]
    Z, Y, X = np.meshgrid(lo[2] + spacing * (kk - org[2]), lo[1] + spacing * (jj - org[1]),
                          lo[0] + spacing * (ii - org[0]), indexing="ij")
    f = _trilinear_fn(c)
    # trilinear in index coordinates = trilinear in world coordinates (axis-aligned map)
    vox = f(X, Y, Z)
    inv = 1.0 / spacing
    w2i = (inv, 0, 0, org[0] - lo[0] * inv, 0, inv, 0, org[1] - lo[1] * inv, 0, 0, inv, org[2] - lo[2] * inv)
    d = _desc(vox, origin=tuple(org), w2i=w2i)
    p = lo + rng.uniform(0, 1, size=(2000, 3)) * spacing * (np.array([nx, ny, nz]) - 1)
    p = p.astype(np.float32)
    got = po.grid_density(d, p).astype(np.float64)
    want = f(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64), p[:, 2].astype(np.float64))
    scale = np.abs(vox).max()
    assert np.max(np.abs(got - want)) < 2e-5 * scale


def test_matches_z_first_double_interpolation_within_reference_tolerance():
    """OpenVDB's BoxSampler interpolates z first, then y, then x; the dense grid
    interpolates x first in float.  Both are the same trilinear interpolant."""
    rng = np.random.default_rng(3)
    vox = rng.uniform(-1, 1, size=(6, 7, 8)).astype(np.float32)
    d = _desc(vox)
    p = (rng.uniform(0, 1, size=(4000, 3)) * np.array([7, 6, 5])).astype(np.float32)
    got = po.grid_density(d, p).astype(np.float64)
    q = p.astype(np.float64)
    i = np.floor(q).astype(int)
    u = q - i
    v = vox.astype(np.float64)

    def at(dx, dy, dz):
        return v[np.minimum(i[:, 2] + dz, 5), np.minimum(i[:, 1] + dy, 6), np.minimum(i[:, 0] + dx, 7)]
    lerp = lambda a, b, t: a + t * (b - a)  # noqa: E731
    vz00 = lerp(at(0, 0, 0), at(0, 0, 1), u[:, 2])
    vz01 = lerp(at(0, 1, 0), at(0, 1, 1), u[:, 2])
    vz10 = lerp(at(1, 0, 0), at(1, 0, 1), u[:, 2])
    vz11 = lerp(at(1, 1, 0), at(1, 1, 1), u[:, 2])
    want = lerp(lerp(vz00, vz01, u[:, 1]), lerp(vz10, vz11, u[:, 1]), u[:, 0])
    err = np.abs(got - want)
    assert err.max() < 2e-4
    assert err.max() < 1e-6  # in fact: float rounding of a few operations


def test_clamp_box_and_voxel_clamp():
    rng = np.random.default_rng(4)
    vox = rng.normal(size=(6, 6, 6)).astype(np.float32)
    # the reference clamps index points to [min + 2, max - 3] of the active bounds
    d = _desc(vox, origin=(10, 10, 10), cmin=(12.0, 12.0, 12.0), cmax=(12.5, 13.0, 12.0))
    p = (rng.uniform(0, 30, size=(500, 3))).astype(np.float32)
    pc = np.stack([np.clip(p[:, 0], 12.0, 12.5), np.clip(p[:, 1], 12.0, 13.0), np.clip(p[:, 2], 12.0, 12.0)], 1)
    assert np.array_equal(po.grid_density(d, p), po.grid_density(d, pc.astype(np.float32)))
    # a clamp box reaching past the lattice: voxel indices clamp to the grid (edge values extend)
    d2 = _desc(vox, origin=(0, 0, 0), cmin=(-5.0, -5.0, -5.0), cmax=(20.0, 20.0, 20.0))
    far = np.array([[-3.0, 2.0, 2.0], [2.0, 2.0, 17.5]], np.float32)
    near = np.array([[0.0, 2.0, 2.0], [2.0, 2.0, 5.0]], np.float32)
    assert np.array_equal(po.grid_density(d2, far), po.grid_density(d2, near))


def test_rotated_world_to_index_map():
    """A world->index map with a rotation: an affine world function tabulated at the
    lattice's world points is reproduced at random world points."""
    rng = np.random.default_rng(5)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    s, t0 = 8.0, np.array([6.0, 6.0, 6.0])
    M = np.concatenate([s * R, t0[:, None]], 1)  # index = s R p + t0
    n = 13
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).astype(np.float64)
    world = np.einsum("ij,...j->...i", np.linalg.inv(s * R), idx - t0)  # [k][j][i] order below
    a = np.array([0.3, -1.2, 0.8])
    vals = world @ a + 0.1  # vals[i][j][k] with idx = (i, j, k) = (x, y, z)
    vox = np.transpose(vals, (2, 1, 0))
    d = _desc(vox, w2i=tuple(M.reshape(-1)))
    p = rng.uniform(-0.3, 0.3, size=(1000, 3)).astype(np.float32)
    got = po.grid_density(d, p).astype(np.float64)
    want = p.astype(np.float64) @ a + 0.1
    assert np.max(np.abs(got - want)) < 1e-5


def test_mean_offset_scale_volume_and_gradient():
    rng = np.random.default_rng(6)
    a = np.array([0.5, -0.25, 2.0])
    n = 12
    Z, Y, X = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    vox = (1.5 + a[0] * X + a[1] * Y + a[2] * Z) / n + 0.2
    d = _desc(vox, offset=0.75, scale=-2.0)
    p = rng.uniform(1, n - 2, size=(300, 3))
    dens = po.grid_density(d, p.astype(np.float32)).astype(np.float64)
    m = po.grid_mean(d, p)
    assert np.allclose(m, (dens + 0.75) * -2.0, rtol=0, atol=1e-12)
    g = po.grid_mean_grad(d, p)
    # the forward difference of an affine function is its slope, up to the float
    # rounding of the shifted lookups (|f| ~ 1, eps = 1e-3: ~1e-4 relative)
    assert np.allclose(g, -2.0 * a / n, rtol=0, atol=2e-3 * np.abs(a).max())
    # the gradient is exactly the difference quotient of the mean at the shifted points
    for k in range(3):
        q = p.copy()
        q[:, k] += 0.001
        assert np.array_equal(g[:, k], (po.grid_mean(d, q) - m) / 0.001)
    # volume: -log(max(1e-4, density)) before offset and scale
    dv = _desc(vox, offset=0.5, scale=3.0, volume=True)
    mv = po.grid_mean(dv, p)
    assert np.allclose(mv, (-np.log(np.maximum(1e-4, dens)) + 0.5) * 3.0, rtol=1e-15, atol=1e-12)
    dz = _desc(np.zeros((3, 3, 3)) - 1.0, volume=True)
    assert np.allclose(po.grid_mean(dz, [[1.0, 1.0, 1.0]]), -np.log(1e-4))


def _sphere_media(js, spacing):
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    med_a = SparseConvolutionNoiseMedium.from_json(js)
    grid = W.sphere_sdf_grid(radius=med_a.mean_radius, center=med_a.mean_center, spacing=spacing)
    med_g = SparseConvolutionNoiseMedium.from_json(W.with_tabulated_mean(js), mean_grid=grid)
    return med_a, med_g


def test_tabulated_sphere_sdf_flights_follow_the_analytic_sphere():
    """3D flights over a finely tabulated sphere SDF (spacing 0.01, interpolation error
    below 1e-4) land where the analytic sphere mean's flights land."""
    from gpis_extend_amd import workloads as W
    from tests.helpers import oracle_field, oracle_rays
    med_a, med_g = _sphere_media(W.config0_json(), 0.01)
    rays = W.sphere_rays(med_a, n=2000, radius=1.0)  # origins inside the tabulated box
    ra = oracle_field(med_a).trace(oracle_rays(rays, med_a), nthreads=NTH)
    rg = oracle_field(med_g).trace(oracle_rays(rays, med_g), nthreads=NTH)
    same = ra["outcome"] == rg["outcome"]
    assert same.mean() > 0.99
    hit = same & (ra["outcome"] == po.HIT)
    assert hit.sum() > 500
    dt = np.abs(ra["t"][hit] - rg["t"][hit])
    assert np.median(dt) < 1e-4 and np.quantile(dt, 0.95) < 2e-3
    # and the tabulated mean itself is the sphere SDF to the interpolation error
    desc = po.make_desc(mean_kind=po.MEAN_GRID, grid=med_g.mean_grid)
    p = np.random.default_rng(7).uniform(-0.9, 0.9, size=(500, 3))
    sdf = np.linalg.norm(p, axis=1) - med_a.mean_radius
    r = np.linalg.norm(p, axis=1)
    # trilinear interpolation error of |p|: at most (h^2 / 8) * (sum of the second
    # derivatives along the axes) <= 3 h^2 / (8 (r - h)), h = 0.01
    ok = r > 0.05
    bound = 3 * 0.01 ** 2 / (8 * (r[ok] - 0.01)) + 1e-6
    assert np.all(np.abs(po.grid_mean(desc, p)[ok] - sdf[ok]) < bound)


def test_tabulated_sphere_sdf_1d_flights_follow_the_analytic_sphere():
    from gpis_extend_amd import workloads as W
    from tests.helpers import oracle_desc, oracle_rays
    med_a, med_g = _sphere_media(W.config3_json(), 0.01)
    rays = W.renewal1d_rays(med_a, n=2000)
    ls = rays["length_scale"]
    ra = po.trace_ls(oracle_desc(med_a), oracle_rays(rays, med_a), ls, nthreads=NTH)
    rg = po.trace_ls(oracle_desc(med_g), oracle_rays(rays, med_g), ls, nthreads=NTH)
    same = ra["outcome"] == rg["outcome"]
    assert same.mean() > 0.99
    hit = same & (ra["outcome"] == po.HIT)
    assert hit.sum() > 300
    dt = np.abs(ra["t"][hit] - rg["t"][hit])
    assert np.median(dt) < 1e-4


def test_medium_json_tabulated_mean():
    from gpis_extend_amd import MediumConfigError, SparseConvolutionNoiseMedium, workloads as W
    js = W.with_tabulated_mean(W.config0_json(), offset=0.25, scale=2.0, volume=True)
    with pytest.raises(MediumConfigError, match="dense export"):
        SparseConvolutionNoiseMedium.from_json(js)
    g = W.sphere_sdf_grid(spacing=0.1)
    m = SparseConvolutionNoiseMedium.from_json(js, mean_grid=g)
    assert m.mean_kind == po.MEAN_GRID and m.mean_offset == 0.25 and m.mean_scale == 2.0 and m.mean_volume
    with pytest.raises(MediumConfigError):
        W.DenseMeanGrid(np.zeros((1, 4, 4)))
