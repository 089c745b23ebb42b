"""The oracle's procedural nonstationary covariance (ProceduralNonstationaryCovariance,
GPFunctions.cpp:1225-1345, 1600-1760; the level ladder of SparseConvolutionNoise.cpp:101-111,
142-153) pinned by the laws its two realizations must follow where the length-scale field
is constant, l(x) = c:

* brute force (multiResolutionGrid false): the splats use the local scale l(p)/l_max on a
  lattice of the largest scale, so the field is the stationary SE field of length c:
  Var = 1 and E[psi(x) psi(x + d)] = exp(-|d|^2 / (2 c^2));
* multi-resolution: two independent levels of the base-2.5 ladder s_lo <= c < s_hi
  blended with w_lo = (1 - r)/sqrt((1 - r)^2 + r^2), w_hi = r/sqrt(...), r = (c - s_lo) /
  (s_hi - s_lo): Var = w_lo^2 + w_hi^2 = 1 and the covariance is the mixture
  w_lo^2 exp(-d^2 / (2 s_lo^2)) + w_hi^2 exp(-d^2 / (2 s_hi^2));
* with the spatially varying field of the nonstationary workload the variance stays 1
  everywhere (binned by the local length scale).

Monte-Carlo z-tests over many independent points; a wrong level weight, level scale or
normalisation fails them.  The per-level seed offset (floor(log s / log 2.5)) is invisible
to these laws (levels of different cell sizes decorrelate whatever their seeds): it is
pinned only by the level-index check of test_oracle_pinning.py."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

N = 150000


def _ns_desc(base, amp, multires, ls_max):
    return po.make_desc(cov_kind=po.COV_PROC_NS, ns_multires=int(multires), ns_ls_field=po.LS_SINCOS,
                        ns_ls_params=(base, amp, 4 * math.pi, 0.0), ns_ls_max=ls_max, length_scale=1.0,
                        sigma=1.0, impulse_density=8, single_realization=1, global_seed=5, ctxt=po.CTX_GLOBAL,
                        mean_kind=po.MEAN_PLANE, plane_normal=(0.0, 0.0, 0.0), plane_offset=0.0)


def _values(F, x, seed):
    return F.eval_points(x.astype(np.float32), seed, nthreads=8)[:, 0].astype(np.float64)


def _cov_check(F, law, lags, rng, box=3.0):
    x = rng.uniform(-box, box, (N, 3))
    v = _values(F, x, 5)
    assert abs(v.var() - 1.0) < 6 * math.sqrt(2.0 / N) + 0.005, v.var()
    for lag in lags:
        u = rng.normal(size=(N, 3))
        u = u / np.linalg.norm(u, axis=1, keepdims=True) * lag
        v2 = _values(F, x + u, 5)
        rho = float(np.mean(v * v2))
        want = law(lag)
        se = math.sqrt((1 + want * want) / N)
        # 0.008: the truncated splats' own departure from the SE law (measured <= 0.009 at
        # 2e5 samples for the stationary field, the largest at lag 1.5 l)
        assert abs(rho - want) < 5 * se + 0.008, (lag, rho, want)


def test_brute_force_constant_scale_is_the_se_field():
    c = 0.1
    F = po.Field(_ns_desc(c, 0.0, False, c))
    _cov_check(F, lambda d: math.exp(-d * d / (2 * c * c)), (0.05, 0.1, 0.15), np.random.default_rng(1))


@pytest.mark.parametrize("c", [0.1, 0.07])
def test_multires_constant_scale_is_the_two_level_mixture(c):
    lvl = math.floor(math.log(c) / math.log(2.5))
    slo, shi = 2.5 ** lvl, 2.5 ** (lvl + 1)
    r = (c - slo) / (shi - slo)
    den = math.sqrt((1 - r) ** 2 + r * r)
    wlo, whi = (1 - r) / den, r / den

    def law(d):
        return wlo * wlo * math.exp(-d * d / (2 * slo * slo)) + whi * whi * math.exp(-d * d / (2 * shi * shi))
    # the mixture differs from SE(c) by more than the test resolves at these lags
    assert abs(law(0.1) - math.exp(-0.01 / (2 * c * c))) > 0.04
    F = po.Field(_ns_desc(c, 0.0, True, c))
    _cov_check(F, law, (0.05, 0.1, 0.15), np.random.default_rng(2))


@pytest.mark.parametrize("multires", [True, False])
def test_unit_variance_under_a_varying_length_scale(multires):
    # l(x) = 0.02 + 0.18 (0.5 + 0.5 sin(4 pi x0) cos(4 pi x2)), the nonstationary workload's field
    F = po.Field(_ns_desc(0.02, 0.18, multires, 0.2))
    rng = np.random.default_rng(3)
    n = 20000 if multires else 6000
    x = rng.uniform(-1, 1, (n, 3))
    v = _values(F, x, 5)
    ls = 0.02 + 0.18 * (0.5 + 0.5 * np.sin(4 * np.pi * x[:, 0]) * np.cos(4 * np.pi * x[:, 2]))
    # points closer than the correlation length are dependent: bin by scale and use a
    # loose bound on the effective sample size
    for lo, hi in ((0.02, 0.08), (0.08, 0.14), (0.14, 0.2)):
        sel = (ls >= lo) & (ls < hi)
        m = int(sel.sum())
        assert m > 500
        assert abs(v[sel].var() - 1.0) < 8 * math.sqrt(2.0 / m) + 0.04, (lo, hi, v[sel].var())


@pytest.mark.parametrize("lvl", [-3, -2])
def test_multires_levels_are_independent(lvl):
    """A constant length scale at a rung of the ladder, c = 2.5^L, gives weights (1, 0): the
    field is level L alone; c = 2.5^(L+1) gives level L+1 alone, with that level's own seed
    offset.  Both media share the global seed, so they expose the two levels of one
    realization: their cross-covariance at equal and nearby points must vanish (z-test)."""
    c_lo = float(np.float32(2.5) ** np.float32(lvl))
    c_hi = float(np.float32(2.5) ** np.float32(lvl + 1))
    A = po.Field(_ns_desc(c_lo, 0.0, True, c_lo))
    B = po.Field(_ns_desc(c_hi, 0.0, True, c_hi))
    rng = np.random.default_rng(11 + lvl)
    # a box of 80 coarse length scales: the samples are nearly independent
    x = rng.uniform(-40.0 * c_hi, 40.0 * c_hi, (N, 3))
    va = _values(A, x, 5)
    vb = _values(B, x, 5)
    # each is a unit-variance level on its own
    assert abs(va.var() - 1.0) < 0.03 and abs(vb.var() - 1.0) < 0.03
    for lag in (0.0, 0.5 * c_lo, c_lo):
        u = rng.normal(size=(N, 3))
        u = u / np.linalg.norm(u, axis=1, keepdims=True) * lag
        vb2 = _values(B, x + u, 5) if lag > 0 else vb
        rho = float(np.mean(va * vb2))
        z = rho * math.sqrt(N)
        print(f"level {lvl} x level {lvl + 1}, lag {lag:.3f}: covariance {rho:+.4f} (z = {z:+.2f})")
        assert abs(z) < 5.0, (lag, rho)
