"""The oracle's correlated transverse gradient law of 1D Renewal+ SE media with
1D_gradient_correlationXY (seTransverseGradientMoments, SparseConvolutionNoise.cpp:78-99,
its use in evaluateGradientNoise1D :186-196, neePDF :742-752 and samplingScheme :68-76),
pinned by the checks of the reference's own test tests/SparseTransverseGradientTest.cpp:
the unconditioned moments, the delta at the origin (the observed gradient is recovered
and the sampling scheme falls back to UNI), the conditional variance and mean along the
ray in closed form, the same variance from an independent solve of the value + 3-gradient
GP block, the empirical mean and variance of sampled gradients, and the NEE density of a
sampled normal against the sampled conditional Gaussian."""
import numpy as np
import pytest

from oracle import pyoracle as po

CASES = ((0.04, 0.05), (0.02, 0.2), (0.08, 0.4))
DIRS = (np.array([0.0, 0.0, 1.0], np.float32),
        (np.array([1.0, 2.0, 3.0]) / np.linalg.norm([1.0, 2.0, 3.0])).astype(np.float32))


def _field(sigma, ell):
    d = po.make_desc(cov_kind=po.COV_SE, sigma=sigma, length_scale=ell, kernel_scale=3.0,
                     mean_kind=po.MEAN_PLANE, plane_normal=(0.0, 0.0, 0.0), plane_offset=0.0,
                     global_seed=17429, impulse_density=10.0, single_realization=0, ctxt=po.CTX_RENEWAL_PLUS,
                     one_d=1, scheme=po.SCHEME_MIS, correlation_xy=1)
    return po.Field(d)


def _frame(rd):
    return po.tangent_frame(rd)  # rows: tangent, bitangent, normal


def _ray(rd, g0, pss=(3, 7, 1, 2), first=0):
    r = np.zeros(1, po.RAY_DTYPE)[0]
    r["origin"] = (0.0, 0.0, 0.0)
    r["dir"] = rd
    r["near_t"], r["far_t"], r["jitter"] = 0.0, 1.0, 0.5
    r["pss"] = pss
    r["scene_seed"] = 29
    r["first_scatter"] = first
    r["last_val"] = 0.0
    r["last_aniso"] = np.asarray(g0, np.float64)
    return r


def _g0(F):
    return (0.3 * F[0] + (-0.2) * F[1] + (-0.8) * F[2]).astype(np.float32)


def _dense_conditional_variance(sigma, ell, p, F):
    """Var of the tangent derivative at p given f, grad f at the origin, for
    k(x, y) = sigma^2 exp(-|x - y|^2 / (2 ell^2)), by a direct solve."""
    def k(d):
        return sigma ** 2 * np.exp(-(d @ d) / (2 * ell ** 2))

    def d1(d, a):      # Cov(d_a f(x), f(y)), d = x - y
        return -(d @ a) / ell ** 2 * k(d)

    def d2(d, a, b):   # Cov(d_a f(x), d_b f(y))
        return ((a @ b) / ell ** 2 - (d @ a) * (d @ b) / ell ** 4) * k(d)

    o = np.zeros(3)
    axes = [F[0].astype(np.float64), F[1].astype(np.float64), F[2].astype(np.float64)]
    K = np.zeros((4, 4))
    K[0, 0] = k(o)
    for i, a in enumerate(axes):
        K[0, i + 1] = K[i + 1, 0] = -d1(o, a)
        for j, b in enumerate(axes):
            K[i + 1, j + 1] = d2(o, a, b)
    t = axes[0]
    d = np.asarray(p, np.float64)
    cross = np.array([d1(d, t)] + [d2(d, t, b) for b in axes])
    prior = d2(o, t, t)
    return prior - cross @ np.linalg.solve(K, cross)


@pytest.mark.parametrize("sigma,ell", CASES)
def test_moments_origin_delta_and_dense_gp(sigma, ell):
    F = None
    for rd in DIRS:
        fld = _field(sigma, ell)
        F = _frame(rd)
        g0 = _g0(F)
        # unconditioned (a first-scatter flight): mean 0, variance 1/2
        ok, mu, var, _ = fld.transverse_moments(_ray(rd, g0, first=1), np.zeros(3, np.float32))
        assert ok and np.all(mu == 0.0) and var == 0.5
        ray = _ray(rd, g0)
        # the observation at the origin is a delta: the scheme falls back to UNI and the
        # gradient of the conditioned realization there is the observed one
        ok, mu, var, scheme = fld.transverse_moments(ray, np.zeros(3, np.float32))
        assert ok and scheme == po.SCHEME_UNI and var == 0.0
        out = fld.conditioned_eval(ray, np.zeros(3, np.float32), 0.0)
        assert np.linalg.norm(out[1:] - g0) < 2e-6
        ia = np.sqrt(2.0) / ell
        for q in (0.0, 1e-7, 0.1, 0.7, 1.0, 3.0, 12.0):
            p = (rd * np.float32(q * ell)).astype(np.float32)
            ok, mu, var, _ = fld.transverse_moments(ray, p)
            rho = np.exp(-0.5 * q * q)
            assert abs(var - 0.5 * (-np.expm1(-q * q))) <= 3e-7
            actual = (mu[0] * F[0] + mu[1] * F[1]) * ia * sigma
            expected = (0.3 * F[0] - 0.2 * F[1]) * rho
            assert np.linalg.norm(actual - expected) < 1e-6
            dense = _dense_conditional_variance(sigma, ell, p, F)
            assert abs(var * 2 * sigma ** 2 / ell ** 2 - dense) <= 3e-6 * max(1.0, dense)


def test_sampled_gradients_and_nee_density():
    """Gradients drawn by the conditioned realization at t = 0.7 ell have the conditional
    mean and variance, and neePDF of a sampled normal equals the conditional Gaussian's
    density carried to normals (density |g|^2 / (4 cos^2))."""
    trials = 3000
    pdfs = 0
    for sigma, ell in CASES:
        for rd in DIRS:
            fld = _field(sigma, ell)
            F = _frame(rd)
            g0 = _g0(F)
            t = np.float32(0.7 * ell)
            p = (rd * t).astype(np.float32)
            _, _, var, _ = fld.transverse_moments(_ray(rd, g0), p)
            v = var * 2 * sigma ** 2 / ell ** 2
            scale = sigma * np.sqrt(2.0) / ell  # amplitude * localToWorld^T of the SE kernel
            expected = np.array([np.exp(-0.245) * 0.3, np.exp(-0.245) * -0.2])
            res = np.zeros((trials, 2))
            for i in range(trials):
                ray = _ray(rd, g0, pss=(3, 7, i, 2))
                g = fld.conditioned_eval(ray, p, float(t))[1:].astype(np.float64)
                local = F.astype(np.float64) @ g
                res[i] = local[:2] - expected
                if i < 20:
                    # the draw is the first Box-Muller pair of the normal seed
                    # xxhash32(pss) + xxhash32(sceneSeed) + 1 (:187-189)
                    L = po.lib()
                    ns = (L.orc_xxhash32_4(3, 7, i, 2) + L.orc_xxhash32_1(29) + 1) & 0xFFFFFFFF
                    z = po.standard_normals(ns, 1)
                    _, mu, var_i, _ = fld.transverse_moments(ray, p)
                    want = mu + np.sqrt(var_i) * np.asarray(z[:2])
                    assert np.allclose(local[:2] / scale, want, rtol=0, atol=2e-6 * max(1.0, np.abs(want).max()))
                if i < 200 and local[2] < -0.02:
                    n = (g / np.linalg.norm(g)).astype(np.float32)
                    cosine = abs(float(n.astype(np.float64) @ rd.astype(np.float64)))
                    density = np.exp(-(res[i] @ res[i]) / (2 * v)) / (2 * np.pi * v)
                    expected_pdf = density * (g @ g) / (4 * cosine * cosine)
                    _, pdf = fld.nee(ray, n, p, float(t))
                    assert abs(pdf - expected_pdf) <= 2e-4 * max(1.0, expected_pdf)
                    pdfs += 1
            for axis in range(2):
                assert abs(res[:, axis].mean()) <= 4.0 * np.sqrt(v / trials)
                assert abs((res[:, axis] ** 2).mean() / v - 1.0) <= 5.0 * np.sqrt(2.0 / trials)
    assert pdfs > 100
