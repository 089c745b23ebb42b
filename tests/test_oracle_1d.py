"""Pins for the oracle's 1D along-ray path (1D_sampling media) and its NEE, from
what the mathematics fixes:

* the 1D sparse-convolution process along a ray has unit variance and the SE
  correlation exp(-ds^2 / (2 l^2)) in the ray parameter s;
* derivative laws: the longitudinal derivative (from the 1D splats) and the two
  transverse components (Box-Muller draws) each have the SE gradient variance
  1 / l^2 per world axis, i.e. 1/2 in isotropic ray units;
* the gradient's component along the ray is the derivative of the value
  (central differences), and at the origin the conditioned realization
  reproduces the observed value and directional derivative (the reference's own
  sanity check, SparseConvolutionNoise.cpp:630-656);
* brute force: re-evaluating the conditioned field on the jittered lattice finds
  the first sign change after the first step (the secondary-ray rule) in the step
  the oracle reports;
* neePDF is the density, over outgoing mirror directions, of the reflection of
  the ray about the gradient drawn from the transverse law (Monte Carlo over the
  law against quadrature of the pdf over the sphere), restricted to gradients
  facing the ray; neeGrad lies on the plane of gradients that law can produce.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po


def desc1d(l=0.1, ctxt=po.CTX_RENEWAL_PLUS, **kw):
    return po.make_desc(sigma=1.0, length_scale=l, impulse_density=8, single_realization=0, global_seed=3,
                        mean_radius=0.5, ctxt=ctxt, one_d=1, scheme=po.SCHEME_MIS, **kw)


def rays1d(n, rng, first=1, far=2.0, jitter=None):
    r = np.zeros(n, po.RAY_DTYPE)
    o = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["origin"] = o
    r["dir"] = d.astype(np.float32)
    r["near_t"] = 0.0
    r["far_t"] = far
    r["jitter"] = rng.uniform(size=n) if jitter is None else jitter
    r["pss"][:, 0] = np.arange(n)
    r["pss"][:, 1] = 77
    r["scene_seed"] = 5
    r["first_scatter"] = first
    return r


def values(desc, rays, q, tseg=1.0):
    n = rays.shape[0]
    return po.conditioned_eval_ls(desc, rays, None, q, np.full(n, tseg, np.float32))


def iso_ray_components(g, d, l, amp=1.0):
    """Gradient of the noise in isotropic ray coordinates: world gradient g = amp *
    (1/lc) * (T x + B y + N z) for a unit-aniso SE field with lc = l / sqrt(2)."""
    lc = l / math.sqrt(2.0)
    out = np.zeros_like(g, dtype=np.float64)
    for i in range(g.shape[0]):
        T, B, N = po.tangent_frame(d[i].astype(np.float32)).astype(np.float64)
        w = g[i].astype(np.float64) * lc / amp
        out[i] = (w @ T, w @ B, w @ N)
    return out


def test_1d_unit_variance_and_se_correlation_along_ray():
    l = 0.1
    D = desc1d(l)
    rng = np.random.default_rng(21)
    n = 40000
    r = rays1d(n, rng)
    o = r["origin"].astype(np.float64)
    d = r["dir"].astype(np.float64)
    s = rng.uniform(0.0, 1.0, (n, 1))
    q1 = (o + s * d).astype(np.float32)
    m1 = np.linalg.norm(q1.astype(np.float64), axis=1) - 0.5
    v1 = values(D, r, q1)[:, 0] - m1
    # each ray is its own realization (per-ray seed), so samples are independent
    assert abs(v1.var() - 1.0) < 6 * math.sqrt(2.0 / n) + 0.01, v1.var()
    assert abs(v1.mean()) < 6 / math.sqrt(n)
    for lag in (0.05, 0.1, 0.2):
        q2 = (o + (s + lag) * d).astype(np.float32)
        m2 = np.linalg.norm(q2.astype(np.float64), axis=1) - 0.5
        v2 = values(D, r, q2)[:, 0] - m2
        rho = float(np.mean(v1 * v2))
        want = math.exp(-lag * lag / (2 * l * l))
        se = math.sqrt((1 + want * want) / n)
        assert abs(rho - want) < 6 * se + 0.015, (lag, rho, want)


def test_1d_gradient_laws_and_longitudinal_difference():
    l = 0.1
    D = desc1d(l)
    rng = np.random.default_rng(22)
    n = 20000
    r = rays1d(n, rng)
    o = r["origin"].astype(np.float64)
    d = r["dir"].astype(np.float64)
    q = (o + 0.5 * d).astype(np.float32)
    out = values(D, r, q)
    mg = q.astype(np.float64) / np.linalg.norm(q.astype(np.float64), axis=1, keepdims=True)
    gi = iso_ray_components(out[:, 1:].astype(np.float64) - mg, r["dir"], l)
    # transverse (Box-Muller / sqrt 2) and longitudinal (1D splat derivative) laws
    for ax in range(3):
        v = gi[:, ax]
        assert abs(v.mean()) < 6 * math.sqrt(0.5 / n), (ax, v.mean())
        assert abs(v.var() - 0.5) < 6 * 0.5 * math.sqrt(2.0 / n) + 0.01, (ax, v.var())
    # the endpoint (t_segment > 0) pair is independent of the origin pair (t_segment = 0)
    out0 = values(D, r, q, tseg=0.0)
    g0 = iso_ray_components(out0[:, 1:].astype(np.float64) - mg, r["dir"], l)
    assert abs(np.corrcoef(g0[:, 0], gi[:, 0])[0, 1]) < 6 / math.sqrt(n)
    assert np.allclose(g0[:, 2], gi[:, 2], atol=1e-5)  # the longitudinal part does not depend on the pair
    # along the ray, the gradient is the derivative of the value
    h = 2e-5
    k = 2000
    qp = (o[:k] + (0.5 + h) * d[:k]).astype(np.float32)
    qm = (o[:k] + (0.5 - h) * d[:k]).astype(np.float32)
    fp = values(D, r[:k], qp)[:, 0].astype(np.float64)
    fm = values(D, r[:k], qm)[:, 0].astype(np.float64)
    fd = (fp - fm) / (2 * h)
    gd = (out[:k, 1:].astype(np.float64) * d[:k]).sum(1)
    err = np.abs(fd - gd) / max(1.0, float(np.median(np.abs(gd))))
    assert np.median(err) < 2e-3, np.median(err)
    assert np.quantile(err, 0.9) < 1e-2, np.quantile(err, 0.9)


@pytest.mark.parametrize("ctxt", [po.CTX_RENEWAL, po.CTX_RENEWAL_PLUS])
@pytest.mark.parametrize("l", [0.02, 0.1, 0.2])
def test_1d_conditioning_reproduces_observation(ctxt, l):
    D = desc1d(l, ctxt=ctxt)
    rng = np.random.default_rng(23)
    n = 64
    r = rays1d(n, rng, first=0)
    r["last_val"] = rng.normal(0.3, 0.2, n)
    gd = rng.normal(0.0, 1.0, n)
    r["last_aniso"] = gd[:, None] * r["dir"].astype(np.float64) + rng.normal(0, 0.5, (n, 3))
    out = values(D, r, r["origin"], tseg=1.0)
    assert np.max(np.abs(out[:, 0] - r["last_val"])) < 1e-4
    if ctxt == po.CTX_RENEWAL_PLUS:
        want = (r["last_aniso"].astype(np.float32).astype(np.float64) * r["dir"]).sum(1)
        got = (out[:, 1:].astype(np.float64) * r["dir"]).sum(1)
        assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) < 2e-3


def test_1d_conditioned_crossing_by_lattice_scan():
    """Secondary flights: the first sign change after the first lattice step (the
    sign reference of a conditioned flight) lies in the step the oracle reports."""
    rng = np.random.default_rng(24)
    checked = 0
    for i in range(60):
        l = float(rng.uniform(0.02, 0.2))
        D = desc1d(l)
        r = rays1d(1, rng, first=0, far=1.5)
        r["last_val"] = rng.normal(0.3, 0.2)
        r["last_aniso"] = rng.normal(0, 1) * r["dir"].astype(np.float64)
        res = po.trace_ls(D, r, None)[0]
        if res["outcome"] == po.EXIT or res["t"] == 0.0:
            continue
        far = np.float32(1.5)
        step = min(np.float32(0.01), np.float32(far / np.float32(8)))
        jit = np.float32(r["jitter"][0])
        t0 = np.float32(np.float32(0.0) + step * jit)
        tl = t0 + np.arange(0, 200) * float(step)
        tl = tl[tl < far]
        o, d = r["origin"][0].astype(np.float64), r["dir"][0].astype(np.float64)
        p = (o[None, :] + tl[:, None] * d[None, :]).astype(np.float32)
        rr = np.repeat(r, len(tl))
        f = values(D, rr, p)[:, 0]
        sg = np.where(f < 0, -1, 1)
        ks = np.nonzero(sg[1:] != sg[0])[0]
        assert ks.size, i
        k = int(ks[0]) + 1
        lo, hi = tl[k] - float(step), tl[k]
        assert lo - 1e-9 <= res["t"] <= hi + 1e-9, (i, lo, hi, res["t"])
        checked += 1
    assert checked >= 25


def _nee_setup(l=0.1, seed=25):
    D = desc1d(l)
    rng = np.random.default_rng(seed)
    while True:
        r = rays1d(1, rng, first=0, far=1.5)
        r["last_val"] = rng.normal(0.3, 0.2)
        r["last_aniso"] = rng.normal(0, 1) * r["dir"].astype(np.float64)
        res = po.trace_ls(D, r, None)[0]
        if res["outcome"] == po.HIT and res["t"] > 0.05:
            break
    o, d = r["origin"][0], r["dir"][0]
    p = (o + np.float32(res["t"]) * d).astype(np.float32)
    return D, r, res, p


def _mirror_frame(D, r, p, l):
    """M = amp * diag(1/lc) [T B N] and the conditioned longitudinal derivative at p."""
    d = r["dir"][0]
    T, B, N = po.tangent_frame(d.astype(np.float32)).astype(np.float64)
    lc = l / math.sqrt(2.0)
    M = np.stack([T, B, N], 1) / lc
    g = values(D, r, p[None, :])[0, 1:].astype(np.float64)
    mg = p.astype(np.float64) / np.linalg.norm(p.astype(np.float64))
    dv = np.linalg.solve(M, g - mg)[2]
    return M, mg, dv


def test_nee_grad_on_gradient_plane():
    l = 0.1
    D, r, res, p = _nee_setup(l)
    M, mg, dv = _mirror_frame(D, r, p, l)
    rng = np.random.default_rng(26)
    n = 64
    h = rng.normal(size=(n, 3))
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    rr = np.repeat(r, n)
    g, pdf = po.nee_ls(D, rr, None, h.astype(np.float32), np.repeat(p[None, :], n, 0), np.full(n, res["t"]))
    for i in range(n):
        z = np.linalg.solve(M, g[i].astype(np.float64) - mg)
        assert abs(z[2] - dv) < 1e-3 * max(1.0, abs(dv)), (z, dv)
        c = np.cross(g[i].astype(np.float64), h[i])
        assert np.linalg.norm(c) < 1e-4 * max(1.0, np.linalg.norm(g[i]))
        # a gradient facing away from h (isect < 0) has no density; facing ones may
        # underflow far out in the Gaussian tail
        if g[i].astype(np.float64) @ h[i] < 0:
            assert pdf[i] == 0.0
        else:
            assert pdf[i] >= 0.0


def test_nee_pdf_is_density_of_mirror_direction():
    l = 0.1
    D, r, res, p = _nee_setup(l, seed=27)
    M, mg, dv = _mirror_frame(D, r, p, l)
    d = r["dir"][0].astype(np.float64)
    rng = np.random.default_rng(28)
    # Monte Carlo over the transverse law: x, y ~ N(0, 1/2)
    ns = 400000
    xy = rng.normal(0.0, math.sqrt(0.5), (ns, 2))
    g = (M @ np.stack([xy[:, 0], xy[:, 1], np.full(ns, dv)], 0)).T + mg
    nrm = g / np.linalg.norm(g, axis=1, keepdims=True)
    facing = (nrm @ d) < 0
    wo = d[None, :] - 2.0 * (nrm @ d)[:, None] * nrm
    # quadrature of the pdf over the sphere (uniform directions)
    nq = 200000
    w = rng.normal(size=(nq, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    hh = w - d[None, :]
    hh /= np.linalg.norm(hh, axis=1, keepdims=True)
    rr = np.repeat(r, nq)
    _, pdf = po.nee_ls(D, rr, None, hh.astype(np.float32), np.repeat(p[None, :], nq, 0), np.full(nq, res["t"]))
    pdf = pdf.astype(np.float64)
    mass_q = 4 * math.pi * pdf.mean()
    mass_mc = facing.mean()
    se_q = 4 * math.pi * pdf.std() / math.sqrt(nq)
    assert abs(mass_q - mass_mc) < 6 * se_q + 0.01, (mass_q, mass_mc)
    # and the direction distribution: first moments and a few caps
    for f in (lambda v: v[:, 0], lambda v: v[:, 1], lambda v: v[:, 2],
              lambda v: (v @ d > 0.5).astype(float), lambda v: (v[:, 1] > 0.0).astype(float)):
        mc = float((f(wo) * facing).mean())
        qv = 4 * math.pi * float((f(w) * pdf).mean())
        se = 4 * math.pi * float((f(w) * pdf).std()) / math.sqrt(nq) + 3 * math.sqrt(0.25 / ns)
        assert abs(mc - qv) < 6 * se + 0.01, (mc, qv)
