"""Pins for the oracle's full anisotropy matrix of the SE covariance ("useAnisoMtx" and
"anisoMtx", GPFunctions.cpp:663-681, 706-719, 723-764, 784-843): with
L = l / sqrt(2) * A the local-to-world matrix, the splat is exp(-x^T (L^-T L^-1) x / 2)
and the field the sparse convolution builds has the covariance

    E[psi(x) psi(x + d)] = exp(-d^T (l^2 A A^T)^-1 d / 2).

* Var[psi] = 1 and the covariance above at lags along and across the matrix's axes
  (Monte-Carlo z-tests); a lag the diagonal of A alone would misjudge is among them;
* the gradient is the gradient of the value (central differences), which fixes the
  splat's first derivative in every coordinate;
* Renewal+ conditioning reproduces the observed value and gradient at the origin, which
  fixes the second derivatives and the inverse of the Hessian (the sanity check of
  SparseConvolutionNoise.cpp:526-531, 568-572);
* a diagonal matrix gives the diagonal "aniso" field to float rounding, and the kernel
  radius is kernelScale * l / sqrt(2) times the largest row sum of A.
The host's parsing of the two JSON fields and its refusals of what it does not serve are
checked here too (no GPU needed).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

L = 0.1


def _rot(ax, ang):
    ax = np.asarray(ax, np.float64)
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


# a rotated ellipsoid (axes 1.6, 0.5, 1.0) and a sheared, non-symmetric matrix
MATS = {
    "rotated": (_rot((1, 2, 0.5), 0.7) @ np.diag([1.6, 0.5, 1.0])).astype(np.float32),
    "sheared": np.array([[1.0, 0.6, 0.0], [0.0, 0.8, -0.3], [0.2, 0.0, 1.2]], np.float32),
}


def _field(A, **kw):
    kw.setdefault("ctxt", po.CTX_GLOBAL)
    kw.setdefault("single_realization", 1)
    d = po.make_desc(sigma=1.0, length_scale=L, impulse_density=8, global_seed=1, mean_radius=0.5,
                     aniso_mtx=A, **kw)
    return po.Field(d)


def _noise(F, x):
    out = F.eval_points(x, 1, nthreads=8)
    m = np.linalg.norm(x.astype(np.float64), axis=1) - 0.5
    return out[:, 0] - m, out


@pytest.mark.parametrize("name", sorted(MATS))
def test_variance_and_covariance(name):
    A = MATS[name].astype(np.float64)
    F = _field(MATS[name])
    Sinv = np.linalg.inv(L * L * A @ A.T)
    rng = np.random.default_rng(5)
    n = 60000
    x = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    v, _ = _noise(F, x)
    assert abs(v.var() - 1.0) < 6 * math.sqrt(2.0 / n) + 0.01
    # lags along A's columns (one correlation length each) and along the world axes
    lags = [0.7 * L * A[:, 0], 0.7 * L * A[:, 1], 0.7 * L * A[:, 2],
            np.array([0.08, 0, 0]), np.array([0, 0.08, 0]), np.array([0.05, -0.05, 0.05])]
    for d in lags:
        v2, _ = _noise(F, (x + d).astype(np.float32))
        rho = float(np.mean(v * v2))
        want = math.exp(-0.5 * float(d @ Sinv @ d))
        se = math.sqrt((1 + want * want) / n)
        assert abs(rho - want) < 6 * se + 0.02, (d, rho, want)


def test_covariance_differs_from_the_diagonal():
    """At the first column lag of the rotated matrix the correlation under the full matrix
    and under the matrix's diagonal differ by far more than the test's error bar."""
    A = MATS["rotated"].astype(np.float64)
    d = 0.7 * L * A[:, 0]
    full = math.exp(-0.5 * float(d @ np.linalg.inv(L * L * A @ A.T) @ d))
    diag = math.exp(-0.5 * float(d @ np.linalg.inv(L * L * np.diag(np.diag(A)) ** 2) @ d))
    assert abs(full - diag) > 0.1


@pytest.mark.parametrize("name", sorted(MATS))
def test_gradient_is_gradient_of_value(name):
    F = _field(MATS[name])
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (2000, 3))
    h = 2e-5
    _, base = _noise(F, x.astype(np.float32))
    for ax in range(3):
        e = np.zeros(3); e[ax] = h
        xp, xm = (x + e).astype(np.float32), (x - e).astype(np.float32)
        fp = F.eval_points(xp, 1)[:, 0].astype(np.float64)
        fm = F.eval_points(xm, 1)[:, 0].astype(np.float64)
        fd = (fp - fm) / (xp.astype(np.float64) - xm.astype(np.float64))[:, ax]
        err = np.abs(fd - base[:, 1 + ax]) / max(1.0, float(np.median(np.abs(base[:, 1 + ax]))))
        assert np.median(err) < 2e-3, np.median(err)
        assert np.quantile(err, 0.9) < 1e-2, np.quantile(err, 0.9)


def _ray(o, dvec, last_val, last_grad):
    r = np.zeros(1, po.RAY_DTYPE)
    r["origin"], r["dir"] = o, dvec
    r["near_t"], r["far_t"], r["jitter"] = 0.0, 1.0, 0.5
    r["pss"] = (7, 3, 1, 2)
    r["scene_seed"] = 2
    r["last_val"] = last_val
    r["last_aniso"] = last_grad
    return r[0]


@pytest.mark.parametrize("name", sorted(MATS))
def test_renewal_plus_reproduces_observation(name):
    F = _field(MATS[name], ctxt=po.CTX_RENEWAL_PLUS, single_realization=0)
    rng = np.random.default_rng(12)
    for _ in range(20):
        o = rng.uniform(-0.7, 0.7, 3).astype(np.float32)
        dvec = rng.normal(size=3); dvec /= np.linalg.norm(dvec)
        tv = float(rng.normal(0, 0.3))
        tg = rng.normal(0, 3, 3)
        out = F.conditioned_eval(_ray(o, dvec.astype(np.float32), tv, tg), o, t_segment=1.0)
        assert abs(out[0] - tv) < 1e-4, out
        assert np.max(np.abs(out[1:] - tg.astype(np.float32))) < 1e-3 * max(1, np.abs(tg).max()), (out, tg)


def test_diagonal_matrix_equals_diagonal_aniso():
    a = (1.5, 0.6, 0.9)
    Fm = _field(np.diag(a).astype(np.float32))
    d = po.make_desc(sigma=1.0, length_scale=L, impulse_density=8, single_realization=1, global_seed=1,
                     mean_radius=0.5, ctxt=po.CTX_GLOBAL, aniso=a)
    Fd = po.Field(d)
    assert Fm.kernel_radius == pytest.approx(Fd.kernel_radius, rel=1e-6)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    a1, a2 = Fm.eval_points(x, 1), Fd.eval_points(x, 1)
    # the matrix path inverts A by cofactors: its entries differ from 1/a in the last bit,
    # and impulses at the truncation radius may fall either side of it
    close = np.all(np.abs(a1 - a2) <= 1e-4 * np.maximum(1, np.abs(a2)), axis=1)
    assert close.mean() > 0.995, close.mean()


def test_kernel_radius_is_largest_row_sum():
    A = MATS["sheared"]
    F = _field(A)
    want = 3.0 * (L * math.sqrt(2.0) / 2) * float(A.astype(np.float64).sum(axis=1).max())
    assert F.kernel_radius == pytest.approx(want, rel=1e-6)


def test_host_parses_and_refuses_aniso_mtx():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    from gpis_extend_amd.medium import MediumConfigError
    from tests.helpers import oracle_field

    def js_with(m, **kw):
        js = W.config0_json()
        js["gaussian_process"]["covariance"].update({"useAnisoMtx": True, "anisoMtx": m})
        js.update(kw)
        return js

    A = MATS["sheared"]
    med = SparseConvolutionNoiseMedium.from_json(js_with(A.reshape(9).tolist()))
    assert med.use_aniso_mtx and med.aniso_mtx == tuple(float(x) for x in A.reshape(9))
    assert med.kernel_radius == pytest.approx(oracle_field(med).kernel_radius, rel=1e-7)
    d = med.to_desc()
    assert d.use_aniso_mtx == 1 and list(d.aniso_mtx) == [float(x) for x in A.reshape(9)]
    # JsonPtr::get(Eigen::Matrix3f) takes an array of exactly 9 numbers
    with pytest.raises(MediumConfigError, match="anisoMtx"):
        SparseConvolutionNoiseMedium.from_json(js_with([1.0, 0.0, 0.0]))
    iso = SparseConvolutionNoiseMedium.from_json(js_with(A.reshape(9).tolist(), isotropic_3D_sampling=True))
    assert iso.use_aniso_mtx and iso.to_desc().isotropic_3d == 1
    with pytest.raises(MediumConfigError, match="1D_sampling"):
        SparseConvolutionNoiseMedium.from_json(js_with(A.reshape(9).tolist(), **{"1D_sampling": True}))
    js = W.config2_json()
    js["gaussian_process"]["covariance"]["cov"].update({"useAnisoMtx": True, "anisoMtx": A.reshape(9).tolist()})
    with pytest.raises(MediumConfigError, match="proc_nonstationary"):
        SparseConvolutionNoiseMedium.from_json(js)


# ---- the full matrix in the isotropic ray space (isotropic_3D_sampling) ----
# The ray space of a flight along d is frame(normalize(W d)).toLocal(W p), W = L^-1: the frame
# is a rotation, so for a fixed direction the field keeps the covariance of the world-space
# field above, its gradient is W^T frame.toGlobal of the ray-space one, and conditioning takes
# L^T delta into the ray space (GPFunctions.cpp:634-644, SparseConvolutionNoise.cpp:345-377,
# 539-556).

def _iso_field(A, **kw):
    kw.setdefault("ctxt", po.CTX_GLOBAL)
    kw.setdefault("single_realization", 1)
    kw.setdefault("mean_kind", po.MEAN_PLANE)
    kw.setdefault("plane_normal", (0.0, 0.0, 0.0))
    d = po.make_desc(sigma=1.0, length_scale=L, impulse_density=8, global_seed=1, aniso_mtx=A, iso3d=1, **kw)
    return po.Field(d)


def test_ray_space_keeps_the_covariance_law():
    A = MATS["sheared"].astype(np.float64)
    Sinv = np.linalg.inv(L * L * A @ A.T)
    F = _iso_field(MATS["sheared"])
    d = np.array([0.3, -0.5, 0.8]); d /= np.linalg.norm(d)
    rng = np.random.default_rng(31)
    n = 20000
    x = rng.uniform(-1, 1, (n, 3))
    seeds = rng.integers(1, 2 ** 31, size=n)
    lags = [0.7 * L * A[:, 0], np.array([0.0, 0.08, 0.0]), np.array([0.05, -0.05, 0.05])]
    vals = np.zeros((len(lags) + 1, n))
    for i in range(n):
        pts = np.stack([x[i]] + [x[i] + lg for lg in lags]).astype(np.float32)
        vals[:, i] = F.eval_points_dir(pts, int(seeds[i]), d)[:, 0]
    assert abs(vals[0].var() - 1.0) < 6 * math.sqrt(2.0 / n) + 0.01
    for k, lg in enumerate(lags, start=1):
        rho = float(np.mean(vals[0] * vals[k]))
        want = math.exp(-0.5 * float(lg @ Sinv @ lg))
        se = math.sqrt((1 + want * want) / n)
        assert abs(rho - want) < 6 * se + 0.02, (lg, rho, want)


def test_ray_space_gradient_is_gradient_of_value():
    F = _iso_field(MATS["rotated"])
    rng = np.random.default_rng(32)
    x = rng.uniform(-1, 1, (2000, 3))
    d = np.array([0.6, 0.0, 0.8], np.float32)
    base = F.eval_points_dir(x.astype(np.float32), 1, d, nthreads=8).astype(np.float64)
    # ray-space coordinates are |W| ~ 14-28 times the world ones, so the float rounding of a
    # query moves f more than in world space: a longer difference step averages it down
    h = 1e-4
    for ax in range(3):
        e = np.zeros(3); e[ax] = h
        xp, xm = (x + e).astype(np.float32), (x - e).astype(np.float32)
        fp = F.eval_points_dir(xp, 1, d, nthreads=8)[:, 0].astype(np.float64)
        fm = F.eval_points_dir(xm, 1, d, nthreads=8)[:, 0].astype(np.float64)
        fd = (fp - fm) / (xp.astype(np.float64) - xm.astype(np.float64))[:, ax]
        err = np.abs(fd - base[:, 1 + ax]) / max(1.0, float(np.median(np.abs(base[:, 1 + ax]))))
        assert np.median(err) < 2e-3 and np.quantile(err, 0.9) < 1e-2


@pytest.mark.parametrize("name", sorted(MATS))
def test_ray_space_renewal_plus_reproduces_observation(name):
    F = _iso_field(MATS[name], ctxt=po.CTX_RENEWAL_PLUS, single_realization=0, mean_kind=po.MEAN_SPHERE,
                   mean_radius=0.5)
    rng = np.random.default_rng(33)
    for _ in range(20):
        o = rng.uniform(-0.7, 0.7, 3).astype(np.float32)
        dvec = rng.normal(size=3); dvec /= np.linalg.norm(dvec)
        tv = float(rng.normal(0, 0.3))
        tg = rng.normal(0, 3, 3)
        out = F.conditioned_eval(_ray(o, dvec.astype(np.float32), tv, tg), o, t_segment=1.0)
        assert abs(out[0] - tv) < 1e-4, out
        assert np.max(np.abs(out[1:] - tg.astype(np.float32))) < 1e-3 * max(1, np.abs(tg).max()), (out, tg)


def test_ray_space_diagonal_matrix_equals_diagonal_aniso():
    a = (1.5, 0.6, 0.9)
    Fm = _iso_field(np.diag(a).astype(np.float32))
    d = po.make_desc(sigma=1.0, length_scale=L, impulse_density=8, global_seed=1, aniso=a, iso3d=1,
                     ctxt=po.CTX_GLOBAL, single_realization=1, mean_kind=po.MEAN_PLANE, plane_normal=(0.0, 0.0, 0.0))
    Fd = po.Field(d)
    rng = np.random.default_rng(34)
    x = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    dirn = np.array([0.48, -0.6, 0.64], np.float32)
    a1, a2 = Fm.eval_points_dir(x, 1, dirn, nthreads=8), Fd.eval_points_dir(x, 1, dirn, nthreads=8)
    close = np.all(np.abs(a1 - a2) <= 1e-4 * np.maximum(1, np.abs(a2)), axis=1)
    assert close.mean() > 0.995, close.mean()
