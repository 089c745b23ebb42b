"""Pins for the restatement of a served flight's endpoint (oracle/served_oracle.py), from
the oracle's own conditioned 1D realizations and from closed forms:

* the conditional mean and covariance of (psi, psi') at a lag from the conditioning
  origin (Renewal and Renewal+) equal the sample moments of many independent
  conditioned realizations of the oracle (linear conditioning makes them exact for
  the sparse-convolution process too, not only for a Gaussian one);
* the slopes of the realizations' down-crossings near a lag follow the Rice density
  |y| N(y; mu, s^2) of that law (KS), and the sense of the first lattice point follows
  the conditional law of the value there;
* the rejection sampler draws the size-biased Gaussian (KS against the closed-form
  CDF), and that CDF equals numerical quadrature of the density;
* a dropped term (the b-dependence of the mean, the cross-covariance) or the wrong
  lag units fail the moment test.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle as po
from oracle import served_oracle as so


def _zero_mean_desc(l, ctxt):
    return po.make_desc(sigma=1.0, length_scale=l, impulse_density=8, single_realization=0, global_seed=3,
                        mean_kind=po.MEAN_PLANE, plane_normal=(0.0, 0.0, 0.0), plane_offset=0.0, ctxt=ctxt,
                        one_d=1, scheme=po.SCHEME_MIS)


def _realizations(n, rng, a, b, l, lags, ctxt, plane=None):
    """Values and derivatives along the ray (isotropic ray units) of n independent
    conditioned realizations at the given lags (isotropic units) from the origin."""
    d = np.array([0.36, -0.48, 0.8])
    o = np.array([0.1, 0.2, -0.3], np.float32)
    lc = l / math.sqrt(2.0)
    desc = _zero_mean_desc(l, ctxt) if plane is None else po.make_desc(
        sigma=1.0, length_scale=l, impulse_density=8, single_realization=0, global_seed=3, mean_kind=po.MEAN_PLANE,
        plane_normal=plane[0], plane_offset=plane[1], ctxt=ctxt, one_d=1, scheme=po.SCHEME_MIS)
    m = len(lags)
    r = np.zeros(n * m, po.RAY_DTYPE)
    r["origin"] = o
    r["dir"] = d.astype(np.float32)
    r["near_t"] = 0.0
    r["far_t"] = 10.0
    r["jitter"] = 0.5
    r["pss"][:, 0] = np.repeat(np.arange(n), m)
    r["pss"][:, 1] = 91
    r["scene_seed"] = 4
    r["first_scatter"] = 0
    mean_o = 0.0 if plane is None else float(np.dot(plane[0], o) - plane[1])
    mg = np.zeros(3) if plane is None else np.asarray(plane[0], np.float64)
    r["last_val"] = a + mean_o
    r["last_aniso"] = (b / lc) * d + mg + np.cross(d, [0.0, 0.0, 1.0]) * 0.7  # plus a transverse part
    s = np.tile(np.asarray(lags, np.float64) * lc, n)
    q = (o.astype(np.float64) + s[:, None] * d).astype(np.float32)
    out = po.conditioned_eval_ls(desc, r, None, q, np.ones(n * m, np.float32)).astype(np.float64)
    qd = q.astype(np.float64)
    mean_q = 0.0 if plane is None else qd @ np.asarray(plane[0], np.float64) - plane[1]
    psi = (out[:, 0] - mean_q).reshape(n, m)
    dpsi = (((out[:, 1:] - mg) @ d) * lc).reshape(n, m)
    return psi, dpsi


@pytest.mark.parametrize("ctxt", [po.CTX_RENEWAL, po.CTX_RENEWAL_PLUS])
def test_conditional_moments_equal_the_realizations(ctxt):
    rng = np.random.default_rng(5)
    plus = ctxt == po.CTX_RENEWAL_PLUS
    a, b, l = 0.7, -0.6, 0.1
    lags = [0.4, 1.0, 1.8, 3.0]
    n = 24000
    psi, dpsi = _realizations(n, rng, a, b, l, lags, ctxt)
    for j, lag in enumerate(lags):
        e1, e2, c11, c12, c22 = so.cond_law(lag, a, b, plus)
        x, y = psi[:, j], dpsi[:, j]
        se1, se2 = math.sqrt(c11 / n), math.sqrt(c22 / n)
        assert abs(x.mean() - e1) < 5 * se1 + 2e-3, (lag, x.mean(), e1)
        assert abs(y.mean() - e2) < 5 * se2 + 2e-3, (lag, y.mean(), e2)
        assert abs(x.var() - c11) < 5 * c11 * math.sqrt(2.0 / n) + 5e-3, (lag, x.var(), c11)
        assert abs(y.var() - c22) < 5 * c22 * math.sqrt(2.0 / n) + 5e-3, (lag, y.var(), c22)
        cov = float(np.mean((x - x.mean()) * (y - y.mean())))
        assert abs(cov - c12) < 5 * math.sqrt(c11 * c22 / n) + 5e-3, (lag, cov, c12)


def test_moment_test_fails_planted_mistakes():
    """The same sample moments reject a law without the b-term of the mean, without the
    cross-covariance, or with lags in world rather than isotropic units."""
    rng = np.random.default_rng(6)
    a, b, l = 0.7, -0.6, 0.1
    lag = 1.0
    n = 24000
    psi, dpsi = _realizations(n, rng, a, b, l, [lag], po.CTX_RENEWAL_PLUS)
    x, y = psi[:, 0], dpsi[:, 0]
    e1, e2, c11, c12, c22 = so.cond_law(lag, a, b, True)
    e1_nob, _, _, _, _ = so.cond_law(lag, a, 0.0, True)
    assert abs(x.mean() - e1_nob) > 10 * math.sqrt(c11 / n)
    cov = float(np.mean((x - x.mean()) * (y - y.mean())))
    assert abs(cov - 0.0) > 10 * math.sqrt(c11 * c22 / n)
    e1w, _, _, _, _ = so.cond_law(lag * l / math.sqrt(2.0), a, b, True)
    assert abs(x.mean() - e1w) > 10 * math.sqrt(c11 / n)


def test_crossing_slopes_follow_the_rice_law():
    """Down-crossings of conditioned realizations over a mean with a slope along the
    ray: each slope, through the Rice CDF of the law at its own lag, is uniform (KS),
    and the number of crossings equals the integrated Rice rate."""
    rng = np.random.default_rng(7)
    l = 0.1
    lc = l / math.sqrt(2.0)
    d = np.array([0.36, -0.48, 0.8])
    normal = (0.3 * d).astype(np.float64)  # a plane mean 0.05 + 0.3 s along the ray
    a, b = 0.2, 0.4
    o = np.array([0.1, 0.2, -0.3])
    off = float(normal @ o) - 0.05
    lags = np.arange(0.6, 3.0001, 0.02)
    n = 20000
    psi, dpsi = _realizations(n, rng, a, b, l, lags, po.CTX_RENEWAL_PLUS, plane=(tuple(normal), off))
    m_s = float(normal @ d)
    s = lags * lc
    f = 0.05 + m_s * s + psi
    fprime = m_s + dpsi / lc  # world derivative along the ray (sigma 1)
    down = (f[:, :-1] > 0) & (f[:, 1:] <= 0)
    ii, jj = np.nonzero(down)
    w = f[ii, jj] / (f[ii, jj] - f[ii, jj + 1])
    slope = (1 - w) * fprime[ii, jj] + w * fprime[ii, jj + 1]
    lag_c = (1 - w) * lags[jj] + w * lags[jj + 1]
    e1, e2, c11, c12, c22 = so.cond_law(lag_c, a, b, True)
    xstar = -(0.05 + m_s * lag_c * lc)
    mu, var = so.slope_law(e1, e2, c11, c12, c22, xstar)
    muY, sY = m_s + mu / lc, np.sqrt(var) / lc
    u = np.sort(so.rice_cdf(-slope / sY, -muY / sY))
    k = len(u)
    ks = float(np.max(np.maximum(np.arange(1, k + 1) / k - u, u - np.arange(k) / k)))
    assert k > 3000, k
    assert ks < 1.63 / math.sqrt(k) + 0.015, (ks, k)
    # planted mistakes fail the same test: no cross-covariance; no |y| weighting
    mu0, var0 = so.slope_law(e1, e2, c11, 0.0 * c12, c22, xstar)
    u0 = np.sort(so.rice_cdf(-slope / (np.sqrt(var0) / lc), -(m_s + mu0 / lc) / (np.sqrt(var0) / lc)))
    assert float(np.max(np.abs(u0 - (np.arange(k) + 0.5) / k))) > 0.05
    from scipy import special
    ug = np.sort(special.ndtr((slope - muY) / sY) / special.ndtr(-muY / sY))
    assert float(np.max(np.abs(ug - (np.arange(k) + 0.5) / k))) > 0.05
    # Rice rate of down-crossings per unit world length: p(f = 0) E[max(-f', 0) | f = 0]
    lm = 0.5 * (lags[:-1] + lags[1:])
    e1, e2, c11, c12, c22 = so.cond_law(lm, a, b, True)
    xs = -(0.05 + m_s * lm * lc)
    mu, var = so.slope_law(e1, e2, c11, c12, c22, xs)
    muY, sY = m_s + mu / lc, np.sqrt(var) / lc
    p0 = np.exp(-0.5 * (xs - e1) ** 2 / c11) / np.sqrt(2 * math.pi * c11)
    z = muY / sY
    e_neg = sY * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi) - muY * 0.5 * np.vectorize(math.erfc)(z / math.sqrt(2))
    rate = float(np.sum(p0 * e_neg * np.diff(s))) * n
    assert abs(k - rate) < 5 * math.sqrt(rate) + 0.03 * rate, (k, rate)


def test_rice_cdf_equals_quadrature_and_sampler_draws_it():
    rng = np.random.default_rng(8)
    for m in (-25.0, -3.0, -0.4, 0.0, 0.6, 2.5, 15.0):
        dens = lambda x: x * math.exp(-0.5 * x * x + m * x - (0.5 * m * m if m > 0 else 0.0))
        hi = max(m, 0.0) + 12.0
        zq = integrate.quad(dens, 0.0, hi, limit=200, points=[max(m, 0.0)])[0]
        for x in np.quantile(so.rice_draw(np.full(4000, m), rng), [0.1, 0.5, 0.9]):
            want = integrate.quad(dens, 0.0, x, limit=200)[0] / zq
            assert abs(so.rice_cdf(x, m) - want) < 1e-7, (m, x)
        xs = np.sort(so.rice_draw(np.full(200000, m), rng))
        cdf = so.rice_cdf(xs, m)
        k = len(xs)
        ks = float(np.max(np.maximum(np.arange(1, k + 1) / k - cdf, cdf - np.arange(k) / k)))
        assert ks < 1.95 / math.sqrt(k), (m, ks)


def test_first_lattice_sign_follows_the_value_law():
    """The sign of the first lattice point's value: the fraction of realizations below zero
    equals Phi((x0 - E1) / sqrt(C11)) of the law."""
    rng = np.random.default_rng(9)
    a, b, l = 0.05, -1.1, 0.1
    lag = 0.5
    n = 30000
    psi, _ = _realizations(n, rng, a, b, l, [lag], po.CTX_RENEWAL_PLUS)
    e1, _, c11, _, _ = so.cond_law(lag, a, b, True)
    p_neg = 0.5 * math.erfc((e1 / math.sqrt(c11)) / math.sqrt(2))
    got = float(np.mean(psi[:, 0] < 0))
    assert abs(got - p_neg) < 5 * math.sqrt(p_neg * (1 - p_neg) / n) + 0.01, (got, p_neg)
