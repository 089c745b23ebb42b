"""Pins for the surrogate's CPU oracle (oracle/mlp_oracle.py): the analytic backward
equals central differences of the loss, Adam follows its published update, and the
piecewise-linear CDF sampler is the inverse of the class/bin distribution."""
import math

import numpy as np

from oracle import mlp_oracle as M


def test_parameter_count_matches_network_shape():
    assert M.N_PARAMS == 8 * 128 + 128 + 3 * (128 * 128 + 128) + 128 * 32 + 32 == 54816


def test_backward_equals_central_differences():
    rng = np.random.default_rng(1)
    theta = M.init_params(seed=3).astype(np.float64)
    x = rng.normal(size=(12, M.N_IN))
    y = rng.integers(0, M.N_CLASS, size=12)
    g = M.gradient(theta, x, y, exact=True)
    idx = rng.choice(M.N_PARAMS, size=160, replace=False)
    # every layer's weights and biases are represented
    offs = np.cumsum([0] + [i * o + o for i, o in M.LAYERS])
    idx = np.concatenate([idx, offs[:-1], offs[1:] - 1])
    h = 1e-6
    for k in idx:
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (M.loss(tp, x, y) - M.loss(tm, x, y)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_rounding_mode_tracks_exact_gradient():
    rng = np.random.default_rng(2)
    theta = M.init_params(seed=4)
    x = rng.normal(size=(256, M.N_IN)).astype(np.float32)
    y = rng.integers(0, M.N_CLASS, size=256)
    ge = M.gradient(theta, x, y, exact=True)
    gb = M.gradient(theta, x, y, exact=False)
    rel = np.linalg.norm(gb - ge) / np.linalg.norm(ge)
    assert rel < 3e-2, rel


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie rounds to even (mantissa 0)
    assert r[2] == 1.0 + 4 * 2 ** -8       # tie rounds to even (up)
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_adam_matches_published_update():
    g = np.array([0.5, -1.0, 0.0], np.float32)
    th, m, v = M.adam_step(np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32), g, 1, lr=1e-3)
    # first step: m_hat = g, v_hat = g^2, update = lr * sign(g) (up to eps)
    assert np.allclose(th, [-1e-3, 1e-3, 0.0], atol=1e-9)


def test_cdf_sampler_inverts_class_distribution():
    rng = np.random.default_rng(5)
    edges = np.concatenate([[0.0], np.sort(rng.uniform(0.01, 0.99, M.N_BINS - 1)), [1.0]])
    p = rng.dirichlet(np.ones(M.N_CLASS))[None, :].repeat(200000, 0)
    u = rng.uniform(size=200000)
    s, ex = M.cdf_sample(p, edges, u)
    c = M.class_of(s, ex, edges)
    freq = np.bincount(c, minlength=M.N_CLASS) / len(u)
    se = np.sqrt(p[0] * (1 - p[0]) / len(u))
    assert np.all(np.abs(freq - p[0]) < 6 * se + 1e-4)
    # uniform density inside a bin
    b = int(np.argmax(p[0, :M.N_BINS]))
    inb = (c == b)
    w = (s[inb] - edges[b]) / (edges[b + 1] - edges[b])
    assert abs(w.mean() - 0.5) < 6 * math.sqrt(1 / 12 / inb.sum())
