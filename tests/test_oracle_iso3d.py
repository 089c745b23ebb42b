"""The oracle's isotropic-ray-space 3D sampling (isotropic_3D_sampling, with the ray
space the reference hard-wires, SparseConvolutionNoise.cpp:30-33 and :142-165, 345-377)
pinned by what the mathematics fixes:

* along +z the ray frame is the identity (TangentFrame), so the ray-space lattice is the
  world-space one scaled by 1/l: values and gradients equal the world-space field's
  (already pinned) to FP32 rounding;
* along +x the frame is a rotation R: the field is the world-space field at R p and its
  gradient is R^T times the world gradient there (a wrong rotation, a missing
  transpose or a missing 1/l fails it);
* for an anisotropic kernel and an oblique ray, Var = 1 and the covariance is the
  anisotropic SE law exp(-sum (d_i / l_i)^2 / 2) (Monte-Carlo z-tests over realizations);
* the gradient is the gradient of the value (central differences);
* Renewal and Renewal+ conditioning in the ray space reproduce the observed value and
  gradient at the origin."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po


def _desc(iso3d, l=0.1, aniso=(1.0, 1.0, 1.0), **kw):
    base = dict(sigma=1.0, length_scale=l, aniso=aniso, impulse_density=8, single_realization=1, global_seed=1,
                mean_kind=po.MEAN_PLANE, plane_normal=(0.0, 0.0, 0.0), plane_offset=0.0, ctxt=po.CTX_GLOBAL,
                iso3d=iso3d)
    base.update(kw)
    return po.make_desc(**base)


def test_along_z_equals_world_space():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    fi = po.Field(_desc(1))
    fw = po.Field(_desc(0))
    a = fi.eval_points_dir(x, 1, (0.0, 0.0, 1.0), nthreads=8).astype(np.float64)
    b = fw.eval_points_dir(x, 1, (0.0, 0.0, 1.0), nthreads=8).astype(np.float64)
    dv = np.abs(a[:, 0] - b[:, 0])
    dg = np.linalg.norm(a[:, 1:] - b[:, 1:], axis=1) / np.maximum(np.linalg.norm(b[:, 1:], axis=1), 1.0)
    # the lattice coordinate is (p / l_conv) / 3 here and p / (3 l_conv) there: they can
    # differ in the last bit, which moves an impulse across the support radius rarely
    assert np.quantile(dv, 0.999) < 1e-5 and np.median(dv) < 1e-6
    assert np.quantile(dg, 0.999) < 5e-5  # FP32 rounding of two gradient chains


def test_along_x_is_the_rotated_world_field():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    F = po.tangent_frame(np.array([1.0, 0.0, 0.0], np.float32)).astype(np.float64)  # rows t, b, n
    fi = po.Field(_desc(1))
    fw = po.Field(_desc(0))
    a = fi.eval_points_dir(x, 1, (1.0, 0.0, 0.0), nthreads=8).astype(np.float64)
    xr = (x.astype(np.float64) @ F.T).astype(np.float32)
    b = fw.eval_points_dir(xr, 1, (0.0, 0.0, 1.0), nthreads=8).astype(np.float64)
    dv = np.abs(a[:, 0] - b[:, 0])
    gb = b[:, 1:] @ F  # R^T g
    dg = np.linalg.norm(a[:, 1:] - gb, axis=1) / np.maximum(np.linalg.norm(gb, axis=1), 1.0)
    assert np.quantile(dv, 0.999) < 1e-5
    assert np.quantile(dg, 0.999) < 5e-5  # FP32 rounding of two gradient chains


def test_anisotropic_se_law_on_an_oblique_ray():
    l, aniso = 0.1, (1.0, 2.0, 0.5)
    la = np.array([l * a for a in aniso])
    d = np.array([0.3, -0.5, 0.8]); d /= np.linalg.norm(d)
    rng = np.random.default_rng(3)
    n = 40000
    x = rng.uniform(-1, 1, (n, 3))
    # independent realizations: one global seed per sample point
    seeds = rng.integers(1, 2 ** 31, size=n)
    lags = [np.array([0.1, 0.0, 0.0]), np.array([0.0, 0.15, 0.0]), np.array([0.02, 0.03, 0.04])]
    vals = {k: np.zeros(n) for k in range(len(lags) + 1)}
    fi = po.Field(_desc(1, l=l, aniso=aniso))
    # one seed per point: evaluate each point and its lagged partners in the same realization
    for i in range(n):
        pts = np.stack([x[i]] + [x[i] + lg for lg in lags]).astype(np.float32)
        o = fi.eval_points_dir(pts, int(seeds[i]), d)
        for k in range(len(lags) + 1):
            vals[k][i] = o[k, 0]
    v0 = vals[0]
    assert abs(v0.var() - 1.0) < 6 * math.sqrt(2.0 / n) + 0.01
    for k, lg in enumerate(lags, start=1):
        rho = float(np.mean(v0 * vals[k]))
        want = math.exp(-0.5 * float(np.sum((lg / la) ** 2)))
        se = math.sqrt((1 + want * want) / n)
        assert abs(rho - want) < 6 * se + 0.02, (lg, rho, want)


def test_gradient_is_gradient_of_value():
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (2000, 3)).astype(np.float64)
    d = np.array([0.6, 0.0, 0.8], np.float32)
    fi = po.Field(_desc(1, aniso=(1.0, 1.5, 0.75)))
    base = fi.eval_points_dir(x.astype(np.float32), 1, d, nthreads=8).astype(np.float64)
    h = 2e-5
    for ax in range(3):
        e = np.zeros(3); e[ax] = h
        xp, xm = (x + e).astype(np.float32), (x - e).astype(np.float32)
        fp = fi.eval_points_dir(xp, 1, d, nthreads=8)[:, 0].astype(np.float64)
        fm = fi.eval_points_dir(xm, 1, d, nthreads=8)[:, 0].astype(np.float64)
        fd = (fp - fm) / (xp.astype(np.float64) - xm.astype(np.float64))[:, ax]
        err = np.abs(fd - base[:, 1 + ax]) / max(1.0, float(np.median(np.abs(base[:, 1 + ax]))))
        assert np.median(err) < 2e-3 and np.quantile(err, 0.9) < 1e-2


def _ray(origin, direction, last_val, last_grad, pss=(7, 3, 1, 2)):
    r = np.zeros(1, po.RAY_DTYPE)
    r["origin"] = origin
    r["dir"] = direction
    r["near_t"], r["far_t"], r["jitter"] = 0.0, 1.0, 0.5
    r["pss"] = pss
    r["scene_seed"] = 2
    r["first_scatter"] = 0
    r["last_val"] = last_val
    r["last_aniso"] = last_grad
    return r[0]


@pytest.mark.parametrize("ctxt", [po.CTX_RENEWAL, po.CTX_RENEWAL_PLUS])
def test_conditioning_reproduces_observation(ctxt):
    F = po.Field(_desc(1, l=0.1, aniso=(1.0, 1.3, 0.8), single_realization=0, ctxt=ctxt,
                       mean_kind=po.MEAN_SPHERE, mean_radius=0.5))
    rng = np.random.default_rng(11)
    for _ in range(20):
        o = rng.uniform(-0.7, 0.7, 3).astype(np.float32)
        dvec = rng.normal(size=3); dvec /= np.linalg.norm(dvec)
        tv = float(rng.normal(0, 0.3))
        tg = rng.normal(0, 3, 3)
        out = F.conditioned_eval(_ray(o, dvec.astype(np.float32), tv, tg), o, t_segment=1.0)
        assert abs(out[0] - tv) < 1e-4, out
        if ctxt == po.CTX_RENEWAL_PLUS:
            assert np.max(np.abs(out[1:] - tg.astype(np.float32))) < 1e-3 * max(1, np.abs(tg).max())
