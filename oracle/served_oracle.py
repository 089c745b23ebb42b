"""CPU restatement of the endpoint of a flight that the learned sampler serves.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by nothing in the product path.

A served flight has the network's distance t but no realization of the field.  The
plain method still draws a gradient at the endpoint and applies resumeDistance's
rejection rules (GaussianProcessMedium.cpp:300-330): a crossing whose gradient points
along the ray (an up-crossing) fails as "outward", a zero or non-finite gradient fails;
an exit carries the field value at the bound (the next conditioning value,
SparseConvolutionNoiseMedium.cpp:339-346).  The served endpoint takes them from the law
of the conditioned process instead of a realization:

* psi(u) is the unit-variance 1D SE process of the flight in isotropic ray units
  (u = s * |worldToLocal(d)|, correlation exp(-u^2/4), derivative variance 1/2: the
  1D law the oracle's pinning tests hold for evaluateValueNoise1D/GradientNoise1D),
  f(s) = m(s) + sigma * psi(k s) with m the mean along the ray;
* conditioning (conditioning1D, SparseConvolutionNoise.cpp:576-656) observes
  psi(0) = a = (v - m(o)) / sigma and, for Renewal+, psi'(0) = b (the observed
  gradient's component along the ray in the same units);
* the crossing's sense is the sign of f at the first lattice point t0 = near +
  step * jitter (a conditioned flight takes its sign reference there,
  SparseConvolutionNoiseMedium.cpp:304-306), drawn from its conditional law;
* at a crossing f(t) = 0 the derivative along the ray has the Rice density
  |y| N(y; mu, s^2) on the sense's side (y < 0 for a down-crossing, a hit), mu and s
  the conditional law of f'(t) given f(t) = 0 and the observations;
* at the "t = 0" atom (the secant's cut-off, the ray origin) the derivative is the
  conditioned one at the origin (b under Renewal+, N(0, 1/2) under Renewal);
* at an exit the value is the conditional law of f(far) restricted to the sense's
  side, the derivative its conditional law given that value.
The transverse gradient is the plain method's own Box-Muller draw of the same seed.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import special

SQRT2 = math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def one_minus_exp_one_plus(x):
    """1 - exp(-x) (1 + x) without cancellation for small x."""
    x = np.asarray(x, np.float64)
    series = x * x * (0.5 + x * (-1.0 / 3.0 + x * (1.0 / 8.0 + x * (-1.0 / 30.0 + x / 144.0))))
    return np.where(x < 1e-3, series, 1.0 - np.exp(-x) * (1.0 + x))


def cond_law(delta, a, b, plus):
    """Conditional mean and covariance of (psi(delta), psi'(delta)) given psi(0) = a and,
    when `plus`, psi'(0) = b; delta in isotropic ray units.  Returns E1, E2, C11, C12, C22."""
    d = np.asarray(delta, np.float64)
    x = 0.5 * d * d
    r2 = np.exp(-x)  # rho^2
    rho = np.exp(-0.5 * x)
    a = np.asarray(a, np.float64)
    if plus:
        b = np.asarray(b, np.float64)
        e1 = rho * (a + d * b)
        e2 = rho * (-0.5 * d * a + (1.0 - x) * b)
        c11 = one_minus_exp_one_plus(x)
        c22 = 0.5 * (1.0 - r2 * (1.0 - x + x * x))
        c12 = 0.25 * d * d * d * r2
    else:
        e1 = rho * a
        e2 = -0.5 * d * rho * a
        c11 = -np.expm1(-x)
        c22 = 0.5 * (1.0 - x * r2)
        c12 = 0.5 * d * r2
    return e1, e2, c11, c12, c22


def slope_law(e1, e2, c11, c12, c22, xstar):
    """Mean and variance of psi' given psi = xstar under the law (e1, e2, c11, c12, c22)."""
    c11 = np.asarray(c11, np.float64)
    ok = c11 > 1e-300
    q = np.where(ok, c12 / np.where(ok, c11, 1.0), 0.0)
    mu = e2 + q * np.where(ok, xstar - e1, 0.0)
    var = np.maximum(c22 - q * c12, 0.0)
    return mu, var


def rice_mode(m):
    m = np.asarray(m, np.float64)
    s = np.sqrt(m * m + 4.0)
    return np.where(m < 0, 2.0 / (s - m), 0.5 * (m + s))


def rice_draw(m, rng):
    """X >= 0 with density proportional to x exp(-x^2/2 + m x) (the size-biased Gaussian
    N(m, 1) on [0, inf)), by the rejection scheme of the device: a Gamma(2) proposal of
    rate 1/mode accepted with exp(-(x - mode)^2/2) for m < 0, a unit normal proposal
    about m + 1/mode accepted with (x/mode) exp(1 - x/mode) for m >= 0."""
    m = np.atleast_1d(np.asarray(m, np.float64))
    out = np.full(m.shape, np.nan)
    todo = np.ones(m.shape, bool)
    xm = rice_mode(m)
    while todo.any():
        i = np.nonzero(todo)[0]
        mi, xi = m[i], xm[i]
        u1, u2, u3 = rng.uniform(size=(3, len(i)))
        xg = -np.log((1.0 - u1) * (1.0 - u2)) * xi
        zg = rng.normal(size=len(i))
        xn = mi + 1.0 / xi + zg
        neg = mi < 0
        x = np.where(neg, xg, xn)
        acc_g = u3 < np.exp(-0.5 * (xg - xi) ** 2)
        r = xn / xi
        acc_n = (xn > 0) & (u3 < r * np.exp(1.0 - r))
        acc = np.where(neg, acc_g, acc_n)
        out[i[acc]] = x[acc]
        todo[i[acc]] = False
    return out


def _log_tail(z, m):
    """log of T = Z(z) + (m - z) Phi(z), Z(z) = phi(z) + z Phi(z): the survival integral
    int_x^inf t phi(t - m) dt at x = m - z, stable for z << 0 through erfcx."""
    z = np.asarray(z, np.float64)
    m = np.broadcast_to(np.asarray(m, np.float64), z.shape)
    neg = z < 0
    zn = np.where(neg, z, -1.0)
    inner = INV_SQRT_2PI + m * 0.5 * special.erfcx(-zn / SQRT2)
    log_neg = -0.5 * zn * zn + np.log(np.maximum(inner, 1e-300))
    zp = np.where(neg, 0.0, z)
    phi = np.exp(-0.5 * zp * zp) * INV_SQRT_2PI
    Phi = special.ndtr(zp)
    log_pos = np.log(np.maximum(phi + zp * Phi + (m - zp) * Phi, 1e-300))
    return np.where(neg, log_neg, log_pos)


def rice_cdf(x, m):
    """CDF of the density proportional to x exp(-x^2/2 + m x) on x >= 0."""
    x = np.maximum(np.asarray(x, np.float64), 0.0)
    m = np.asarray(m, np.float64)
    return -np.expm1(_log_tail(m - x, m) - _log_tail(m, m))


def truncated_normal_cdf_above(z, alpha):
    """CDF of N(0, 1) restricted to z > alpha."""
    return np.clip((special.ndtr(z) - special.ndtr(alpha)) / special.ndtr(-alpha), 0.0, 1.0)
