"""CPU oracle of the learned flight sampler (test infrastructure only: tests/ and
bench.py's CPU-baseline leg import it; the product path never does).

The surrogate has no counterpart in the reference.  Its specification is the
flight sampler it stands in for: a multilayer perceptron

    8 features -> 128 -> 128 -> 128 -> 128 -> 32 logits, GELU (tanh form) between layers,

whose softmax over 32 classes is a piecewise-linear CDF over the normalised
next-crossing distance s = (t - near) / (far - near): classes 0..30 are distance
bins [edge_b, edge_b+1) with uniform density inside a bin, class 31 is "no crossing
before far" (an exit).  Training minimises the mean cross-entropy of the class of
each plain-method flight, with Adam.

Two forward/backward modes:
* exact=True  -- float64 throughout (the mathematics; finite differences pin it);
* exact=False -- the GPU kernels' rounding points: weights, layer inputs and the
  backward operands rounded to bfloat16 (round to nearest even), products summed
  in higher precision, biases and pre-activations kept in float.
"""
from __future__ import annotations

import math

import numpy as np

N_IN = 8
WIDTH = 128
N_HIDDEN = 4
N_CLASS = 32
N_BINS = N_CLASS - 1
LAYERS = [(N_IN, WIDTH)] + [(WIDTH, WIDTH)] * (N_HIDDEN - 1) + [(WIDTH, N_CLASS)]
N_PARAMS = sum(i * o + o for i, o in LAYERS)  # 54,816

ADAM_B1, ADAM_B2, ADAM_EPS = 0.9, 0.999, 1e-8


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and return them as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


def param_views(theta):
    """Split the flat parameter vector (layer by layer: W[in][out] row-major, then b[out])
    into a list of (W, b) views."""
    out, o = [], 0
    for i, n in LAYERS:
        W = theta[o:o + i * n].reshape(i, n)
        o += i * n
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    return out


def init_params(seed=11):
    """He-uniform weights, zero biases (the GPU's gpis_net_create draws the same law)."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(N_PARAMS, np.float32)
    for (W, b), (i, n) in zip(param_views(theta), LAYERS):
        lim = math.sqrt(6.0 / i)
        W[...] = rng.uniform(-lim, lim, size=(i, n)).astype(np.float32)
    return theta


_K = math.sqrt(2.0 / math.pi)


def _sig2u(z):
    """sigmoid(2u) = 0.5 (1 + tanh(u)), u = sqrt(2/pi) (z + 0.044715 z^3)."""
    z = np.asarray(z, np.float64)
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(-2.0 * _K * (z + 0.044715 * z ** 3)))


def gelu(z):
    """GELU in its tanh form (the network's activation)."""
    return z * _sig2u(z)


def gelu_grad(z):
    s = _sig2u(z)
    return s + z * s * (1.0 - s) * 2.0 * _K * (1.0 + 3.0 * 0.044715 * np.asarray(z, np.float64) ** 2)


def _mm(a, b, exact):
    if exact:
        return a.astype(np.float64) @ b.astype(np.float64)
    return (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32)


def forward(theta, x, exact=True):
    """x: [B, 8] normalised features.  Returns (logits, cache)."""
    layers = param_views(theta)
    h = np.asarray(x, np.float64 if exact else np.float32)
    if not exact:
        h = bf16(h)
    xs, zs = [], []
    for li, (W, b) in enumerate(layers):
        Wq = W.astype(np.float64) if exact else bf16(W)
        xs.append(h)
        z = _mm(h, Wq, exact) + (b.astype(np.float64) if exact else b.astype(np.float32))
        zs.append(z)
        if li + 1 < len(layers):
            h = gelu(z)
            if not exact:
                h = bf16(h.astype(np.float32))
    return zs[-1], (xs, zs)


def softmax(logits):
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    return e / e.sum(axis=1, keepdims=True)


def loss(theta, x, y, exact=True):
    logits, _ = forward(theta, x, exact)
    p = softmax(logits.astype(np.float64))
    return float(-np.mean(np.log(p[np.arange(len(y)), y])))


def gradient(theta, x, y, exact=True):
    """d(mean cross-entropy)/d theta, flat, same layout as theta."""
    layers = param_views(theta)
    logits, (xs, zs) = forward(theta, x, exact)
    B = len(y)
    p = softmax(logits.astype(np.float64))
    dz = p.copy()
    dz[np.arange(B), y] -= 1.0
    if not exact:
        dz = bf16(dz.astype(np.float32))
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        Wq = W.astype(np.float64) if exact else bf16(W)
        dW = _mm(xs[li].T, dz, exact)
        db = dz.astype(np.float64).sum(axis=0)
        grads[li] = (dW / B, db / B)
        if li > 0:
            dy = _mm(dz, Wq.T, exact)
            dz = dy * gelu_grad(zs[li - 1])
            if not exact:
                dz = bf16(dz.astype(np.float32))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return flat


def adam_step(theta, m, v, g, step, lr=1e-3):
    """One Adam update (step counts from 1); returns new (theta, m, v) in float32."""
    m = ADAM_B1 * m + (1 - ADAM_B1) * g
    v = ADAM_B2 * v + (1 - ADAM_B2) * g * g
    mh = m / (1 - ADAM_B1 ** step)
    vh = v / (1 - ADAM_B2 ** step)
    theta = theta - lr * mh / (np.sqrt(vh) + ADAM_EPS)
    return theta.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def class_of(s, exited, edges):
    """Class index of normalised crossing distances s (clamped to [0, 1)), exits -> 31.
    Class 0 holds s <= 0 (an atom at the origin when edges[1] == 0); otherwise the
    class is the last bin whose lower edge is <= s."""
    s = np.clip(np.asarray(s, np.float64), 0.0, np.nextafter(1.0, 0.0))
    e = np.asarray(edges, np.float64)
    c = np.searchsorted(e[1:N_BINS], s, side="right")
    c = np.where(s > 0, c, 0)
    return np.where(exited, N_BINS, c).astype(np.int32)


def cdf_sample(p, edges, u):
    """Inverse of the piecewise-linear CDF: returns (s, exited) for uniforms u."""
    p = np.asarray(p, np.float64)
    c = np.cumsum(p, axis=1)
    cross = c[:, N_BINS - 1]
    exited = u >= cross
    b = np.array([np.searchsorted(c[i, :N_BINS], u[i], side="right") for i in range(len(u))])
    b = np.minimum(b, N_BINS - 1)
    lo = np.where(b > 0, c[np.arange(len(u)), np.maximum(b - 1, 0)], 0.0)
    pb = p[np.arange(len(u)), b]
    frac = np.where(pb > 0, (u - lo) / np.where(pb > 0, pb, 1.0), 0.0)
    frac = np.clip(frac, 0.0, 1.0)
    e = np.asarray(edges, np.float64)
    s = e[b] + frac * (e[b + 1] - e[b])
    return np.where(exited, 1.0, s), exited


def entropy(p):
    p = np.asarray(p, np.float64)
    return -(p * np.log(np.maximum(p, 1e-30))).sum(axis=1)
