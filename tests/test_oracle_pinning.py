"""Pins for the oracle beyond the reference's golden leaf vectors: what the
mathematics fixes about the realization and the flight.

* normalisation: Var[psi] = 1 and E[psi(x) psi(x+d)] = exp(-|d|^2 / (2 l^2)) (the
  SE covariance the sparse convolution approximates; Monte-Carlo z-tests in the
  spirit of tools/surface_compare/lattice_multires_covariance.cpp:203-241);
* the gradient is the gradient of the value (central differences);
* Renewal/Renewal+ conditioning reproduces the observed value and gradient at
  the origin (the reference's own sanity check, SparseConvolutionNoise.cpp:526-531,568-572);
* a planar mean with zero amplitude is hit at the analytic distance with the
  analytic gradient (src/gpu/tests/RaySegmentBatchTest.cpp:85-119);
* brute force: the first sign change of a dense scan of the oracle's own field
  lies inside the lattice step where the oracle reports the crossing;
* the float level formula of the multi-resolution ladder agrees with exact
  arithmetic on every level the configs reach.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po


def se_field(l=0.1, N=8, seed=1, **kw):
    d = po.make_desc(sigma=1.0, length_scale=l, impulse_density=N, single_realization=1, global_seed=seed,
                     mean_radius=0.5, ctxt=po.CTX_GLOBAL, **kw)
    return po.Field(d)


def noise_values(F, x, seed=1):
    out = F.eval_points(x, seed, nthreads=8)
    m = np.linalg.norm(x.astype(np.float64), axis=1) - 0.5
    return out[:, 0] - m, out


def test_unit_variance_and_se_covariance():
    F = se_field(l=0.1)
    rng = np.random.default_rng(3)
    n = 60000
    x = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    v, _ = noise_values(F, x)
    # Var[psi] = 1: standard error of the sample variance ~ sqrt(2/n_eff); points are
    # spread over 60^3 correlation volumes, so they are nearly independent
    assert abs(v.var() - 1.0) < 6 * math.sqrt(2.0 / n) + 0.01
    for lag in (0.05, 0.1, 0.2):
        u = rng.normal(size=(n, 3))
        u = u / np.linalg.norm(u, axis=1, keepdims=True) * lag
        v2, _ = noise_values(F, (x + u).astype(np.float32))
        rho = float(np.mean(v * v2))
        want = math.exp(-lag * lag / (2 * 0.1 * 0.1))
        se = math.sqrt((1 + want * want) / n)
        assert abs(rho - want) < 6 * se + 0.02, (lag, rho, want)


def test_gradient_is_gradient_of_value():
    # The truncated splats make the field jump by ~0.009 where an impulse leaves the
    # support radius (about 25 such boundaries per cell length), so the difference
    # step is kept short (2e-5) and checked by quantiles.
    F = se_field(l=0.1)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (2000, 3)).astype(np.float64)
    h = 2e-5
    _, base = noise_values(F, x.astype(np.float32))
    for ax in range(3):
        e = np.zeros(3); e[ax] = h
        xp, xm = (x + e).astype(np.float32), (x - e).astype(np.float32)
        fp = F.eval_points(xp, 1)[:, 0].astype(np.float64)
        fm = F.eval_points(xm, 1)[:, 0].astype(np.float64)
        fd = (fp - fm) / (xp.astype(np.float64) - xm.astype(np.float64))[:, ax]
        err = np.abs(fd - base[:, 1 + ax]) / max(1.0, float(np.median(np.abs(base[:, 1 + ax]))))
        assert np.median(err) < 2e-3, np.median(err)
        assert np.quantile(err, 0.9) < 1e-2, np.quantile(err, 0.9)


def _ray(origin, direction, first=0, last_val=0.0, last_grad=(0, 0, 0), pss=(7, 3, 1, 2), scene_seed=2,
         near=0.0, far=1.0, jitter=0.5):
    r = np.zeros(1, po.RAY_DTYPE)
    r["origin"] = origin
    r["dir"] = direction
    r["near_t"], r["far_t"], r["jitter"] = near, far, jitter
    r["pss"] = pss
    r["scene_seed"] = scene_seed
    r["first_scatter"] = first
    r["last_val"] = last_val
    r["last_aniso"] = last_grad
    return r[0]


@pytest.mark.parametrize("ctxt", [po.CTX_RENEWAL, po.CTX_RENEWAL_PLUS])
@pytest.mark.parametrize("multires", [False, True])
def test_conditioning_reproduces_observation(ctxt, multires):
    if multires:
        d = po.make_desc(cov_kind=po.COV_PROC_NS, ns_multires=1, ns_ls_field=po.LS_SINCOS,
                         ns_ls_params=(0.02, 0.18, 4 * math.pi, 0.0), ns_ls_max=0.2, length_scale=1.0,
                         impulse_density=8, ctxt=ctxt, mean_kind=po.MEAN_BOX, mean_half=(0.6, 0.6, 0.6))
    else:
        d = po.make_desc(length_scale=0.1, impulse_density=8, ctxt=ctxt, mean_radius=0.5)
    F = po.Field(d)
    rng = np.random.default_rng(11)
    for _ in range(20):
        o = rng.uniform(-0.7, 0.7, 3).astype(np.float32)
        dvec = rng.normal(size=3); dvec /= np.linalg.norm(dvec)
        target_v = float(rng.normal(0, 0.3))
        target_g = rng.normal(0, 3, 3)
        r = _ray(o, dvec.astype(np.float32), first=0, last_val=target_v, last_grad=target_g)
        out = F.conditioned_eval(r, o, t_segment=1.0)
        assert abs(out[0] - target_v) < 1e-4, out
        if ctxt == po.CTX_RENEWAL_PLUS:
            assert np.max(np.abs(out[1:] - target_g.astype(np.float32))) < 1e-3 * max(1, np.abs(target_g).max())


def test_plane_with_zero_amplitude_hits_analytically():
    d = po.make_desc(sigma=0.0, length_scale=0.3, impulse_density=4, single_realization=1, global_seed=3,
                     mean_kind=po.MEAN_PLANE, plane_normal=(-1.0, 0.0, 0.0), plane_offset=-2.5,
                     ctxt=po.CTX_GLOBAL)
    F = po.Field(d)
    rays = np.zeros(4, po.RAY_DTYPE)
    for i, jit in enumerate((0.0, 0.25, 0.5, 0.99)):
        rays[i] = _ray((0, 0.1, 0), (1, 0, 0), first=1, far=5.0, jitter=jit)
    res = F.trace(rays)
    assert np.all(res["outcome"] == po.HIT)
    assert np.max(np.abs(res["t"] - 2.5)) < 1e-5
    assert np.allclose(res["grad"], [[-1, 0, 0]] * 4, atol=2e-3)
    # miss: the far end falls short of the plane; lastVal = f(far) = 2.5 - far
    miss = _ray((0, 0.1, 0), (1, 0, 0), first=1, far=2.0, jitter=0.3)
    r2 = F.trace(np.array([miss]))
    assert r2["outcome"][0] == po.EXIT
    assert abs(r2["t"][0] - 2.0) == 0.0
    assert abs(r2["value"][0] - 0.5) < 1e-6


def test_crossing_inside_reported_step_by_dense_scan():
    """Independent re-evaluation of the jittered lattice t_j = t0 + j*step finds the
    first sign change at the step the oracle reports; a dense scan confirms a root
    inside that bracket; and the reported secant point lies in the bracket."""
    F = se_field(l=0.1)
    rng = np.random.default_rng(5)
    checked = 0
    for i in range(40):
        o = rng.normal(size=3); o = 2.0 * o / np.linalg.norm(o)
        tgt = rng.normal(size=3) * 0.3
        dvec = tgt - o; dvec /= np.linalg.norm(dvec)
        o32, d32 = o.astype(np.float32), dvec.astype(np.float32)
        far = np.float32(-2.0 * np.dot(o, dvec))
        jit = np.float32(rng.uniform())
        r = _ray(o32, d32, first=1, far=far, jitter=jit)
        res = F.trace(np.array([r]))[0]
        if res["outcome"] == po.EXIT or res["t"] == 0.0:
            continue
        step = min(np.float32(0.01), np.float32(far / np.float32(8)))
        t0 = np.float32(np.float32(0.0) + step * jit)
        tl = np.concatenate([[0.0], t0 + np.arange(0, 1000) * float(step)])
        tl = tl[tl < far]
        p = (o32.astype(np.float64)[None, :] + tl[:, None] * d32.astype(np.float64)[None, :]).astype(np.float32)
        f = F.eval_points(p, 1)[:, 0]
        sg = np.where(f < 0, -1, 1)
        k = int(np.nonzero(sg[1:] != sg[0])[0][0]) + 1
        # the secant brackets [t_k - step, t_k]; for k = 1 that starts before nearT
        # (the reference pairs f(nearT) with t0 - step, SparseConvolutionNoiseMedium.cpp:311-316)
        lo, hi = tl[k] - float(step), tl[k]
        assert lo - 1e-9 <= res["t"] <= hi + 1e-9, (i, lo, hi, res["t"])
        ts = np.linspace(max(lo, tl[k - 1]), hi, 2001)
        pd = (o32.astype(np.float64)[None, :] + ts[:, None] * d32.astype(np.float64)[None, :]).astype(np.float32)
        fd = F.eval_points(pd, 1)[:, 0]
        assert np.any(np.where(fd < 0, -1, 1) != sg[0]), i
        checked += 1
    assert checked >= 20


def test_multires_level_index_float_matches_exact():
    # floor(log(2.5^L) / log(2.5)) evaluated in float (as the reference does on float
    # arguments) must give L for every level the nonstationary configs reach
    for L in range(-12, 6):
        s = np.float32(np.power(np.float32(2.5), np.float32(L)))
        lv = math.floor(float(np.float32(np.log(s)) / np.float32(np.log(np.float32(2.5)))))
        assert lv == L, (L, lv)


def test_impulse_statistics():
    # positions uniform in the cell and weights balanced: chi-square / binomial checks
    pts, ws = [], []
    for c in range(400):
        imp = po.cell3d(c, -c, 3 * c, 9, 8)
        pts.append(imp[:, :3]); ws.append(imp[:, 3])
    pts = np.concatenate(pts); ws = np.concatenate(ws)
    assert set(np.unique(ws)) == {-1.0, 1.0}
    assert abs(ws.mean()) < 5 / math.sqrt(len(ws))
    h, _ = np.histogram(pts[:, 0], bins=10, range=(0, 1))
    chi2 = ((h - len(pts) / 10) ** 2 / (len(pts) / 10)).sum()
    assert chi2 < 40
