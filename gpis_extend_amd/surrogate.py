"""The learned flight sampler (include/gpis_surrogate.h) from Python.

A small MLP maps the state of a conditioned 1D flight to a piecewise-linear CDF over
its normalised next-crossing distance; it is trained on flights of the plain GPU
method and sampled with the flight's own jitter.  The plain method stays the reference
(KS distance, image error) and the fallback (out-of-hull, high-entropy and
unconditioned rays).  All compute runs in libgpis_hip.so; torch only holds device
memory and streams and gathers training batches.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _capi as A
from .medium import GpisDevice, RayBatch, SparseConvolutionNoiseMedium

N_FEAT, N_CLASS, N_BINS, N_PARAMS = A.NET_FEATURES, A.NET_CLASSES, A.NET_CLASSES - 1, A.NET_PARAMS
FEATURE_NAMES = ("value", "dir_derivative", "length_scale", "variance", "bound_length", "mean_origin",
                 "mean_slope", "kernel_se")
SEED_TRAIN, SEED_INFER = 11, 12
# Polyak averaging of the served network: the moving average of the Adam iterates with this
# beta (a time constant of 1,000 steps, four epochs of the 2^24-flight training set)
AVERAGE_BETA = 0.999
# rays whose predicted entropy exceeds this quantile of the training set's fall back to the
# plain method
ENTROPY_QUANTILE = 0.999


@dataclass
class Calibration:
    mean: np.ndarray
    inv_std: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    edges: np.ndarray
    entropy_max: float = float("inf")

    def to_c(self) -> A.NetCalibration:
        c = A.NetCalibration()
        for i in range(N_FEAT):
            c.feat_mean[i] = float(self.mean[i])
            c.feat_inv_std[i] = float(self.inv_std[i])
            c.hull_lo[i] = float(self.lo[i])
            c.hull_hi[i] = float(self.hi[i])
        for b in range(N_CLASS):
            c.edges[b] = float(self.edges[b])
        c.entropy_max = float(self.entropy_max) if math.isfinite(self.entropy_max) else 3.4e38
        return c


def minimax_edges(samples: np.ndarray, n_bins: int = N_BINS, grid: int = 20001, strata=None) -> np.ndarray:
    """Bin edges over [0, 1] that minimise the largest gap between an empirical CDF of
    `samples` and its piecewise-linear interpolation through the edges (bins of uniform
    density): bisection on the tolerance, each candidate filled greedily with the longest
    bins that keep the gap within it.  With `strata` (an integer label per sample) the gap
    is the largest over the strata's own CDFs, so the bins are fine wherever any stratum's
    distribution bends: a flight with a short length scale crosses within a few
    hundredths of its bound, where the pooled CDF is nearly straight.  Quantile edges of
    the pooled sample leave one wide tail bin whose linearisation error alone exceeds
    0.01 in KS."""
    x = np.clip(np.asarray(samples, np.float64), 0.0, 1.0)
    lab = np.zeros(len(x), np.int64) if strata is None else np.asarray(strata)
    groups = [np.sort(x[lab == k]) for k in np.unique(lab)]
    pooled = np.sort(x)
    pts = [np.linspace(0.0, 1.0, grid), np.quantile(pooled, np.linspace(0, 1, grid))]
    if strata is not None:
        pts.append(np.geomspace(1e-4, 1.0, grid // 2))
    xs = np.unique(np.concatenate(pts))
    Fr = np.stack([np.searchsorted(g, xs, side="right") / len(g) for g in groups])
    Fl = np.stack([np.searchsorted(g, xs, side="left") / len(g) for g in groups])

    def gap(i, j):
        a, b = xs[i], xs[j]
        model = Fr[:, i:i + 1] + (Fl[:, j:j + 1] - Fr[:, i:i + 1]) * (xs[i:j + 1] - a) / max(b - a, 1e-300)
        return max(np.max(np.abs(Fr[:, i:j] - model[:, :-1])), np.max(np.abs(Fl[:, i + 1:j + 1] - model[:, 1:])))

    def fill(eps):
        idx, i = [0], 0
        while i < len(xs) - 1:
            lo, hi, best = i + 1, len(xs) - 1, i + 1
            while lo <= hi:
                j = (lo + hi) // 2
                if gap(i, j) <= eps:
                    best, lo = j, j + 1
                else:
                    hi = j - 1
            i = best
            idx.append(i)
            if len(idx) > n_bins + 1:
                return None
        return idx

    lo, hi = 1e-6, 0.5
    for _ in range(32):
        mid = (lo * hi) ** 0.5
        if fill(mid) is None:
            lo = mid
        else:
            hi = mid
    idx = fill(hi)
    edges = [xs[k] for k in idx]
    # pad to exactly n_bins bins by splitting the widest bins
    while len(edges) < n_bins + 1:
        w = np.diff(edges)
        k = int(np.argmax(w))
        edges.insert(k + 1, 0.5 * (edges[k] + edges[k + 1]))
    e = np.asarray(edges, np.float64)
    e[0], e[-1] = 0.0, 1.0
    return e


# features whose quantile strata shape the bin edges: value, directional derivative, length scale
EDGE_STRATA_FEATURES = (0, 1, 2)


def calibrate(feats, s, exited, n_sample: int = 1 << 19, strata_bins: int = 3) -> Calibration:
    """Standardisation, hull, the atom class at s = 0 and the bin edges from the training
    features and normalised distances: 30 bins over the continuous part of the crossings,
    minimax over the quantile strata of (value, derivative, length scale)."""
    import torch
    f = feats.double()
    mean = f.mean(0).cpu().numpy()
    std = f.std(0).cpu().numpy()
    inv_std = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-30), 1.0)
    lo = feats.min(0).values.cpu().numpy().astype(np.float64)
    hi = feats.max(0).values.cpu().numpy().astype(np.float64)
    cross = ~exited
    sc, fc = s[cross], feats[cross]
    if sc.numel() > n_sample:
        g = torch.Generator(device=sc.device)
        g.manual_seed(SEED_TRAIN)
        pick = torch.randperm(sc.numel(), device=sc.device, generator=g)[:n_sample]
        sc, fc = sc[pick], fc[pick]
    sv = sc.float().cpu().numpy()
    sid = strata_of(fc[:, list(EDGE_STRATA_FEATURES)].cpu().numpy(), strata_bins)
    pos = sv > 0
    # class 0: the atom at s = 0 (crossings returned at the ray origin); 30 bins after it
    edges = np.concatenate([[0.0], minimax_edges(sv[pos], n_bins=N_BINS - 1, grid=4001, strata=sid[pos])])
    edges = edges.astype(np.float32)
    for b in range(2, N_BINS + 1):  # strictly increasing in float32 after the atom
        if edges[b] <= edges[b - 1]:
            edges[b] = np.nextafter(edges[b - 1], np.float32(2.0))
    edges[-1] = 1.0
    return Calibration(mean.astype(np.float32), inv_std.astype(np.float32), lo.astype(np.float32),
                       hi.astype(np.float32), edges)


class FlightSurrogate:
    """One gpis_net on the current HIP device."""

    def __init__(self, seed: int = SEED_TRAIN):
        import torch
        if not torch.cuda.is_available():
            raise A.GpisError("no HIP device visible")
        self.L = A.lib()
        h = C.c_void_p()
        A.check(self.L.gpis_net_create(seed, C.byref(h)), None, "gpis_net_create")
        self.h = h
        self.calibration: Optional[Calibration] = None
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    def close(self):
        if getattr(self, "h", None):
            self.L.gpis_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        A.check(st, None, what, net=self.h)

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    # ---- parameters and calibration --------------------------------------
    def set_params(self, theta: np.ndarray):
        a = np.ascontiguousarray(theta, np.float32)
        self._check(self.L.gpis_net_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size),
                    "gpis_net_set_params")

    def get_params(self) -> np.ndarray:
        a = np.zeros(N_PARAMS, np.float32)
        self._check(self.L.gpis_net_get_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size),
                    "gpis_net_get_params")
        return a

    def set_calibration(self, cal: Calibration):
        c = cal.to_c()
        self._check(self.L.gpis_net_set_calibration(self.h, C.byref(c)), "gpis_net_set_calibration")
        self.calibration = cal

    # ---- data --------------------------------------------------------------
    def features(self, dev: GpisDevice, rays: RayBatch):
        import torch
        out = torch.empty((rays.n, N_FEAT), dtype=torch.float32, device="cuda")
        soa = rays.soa()
        A.check(self.L.gpis_net_features_1d(dev.h, C.byref(soa), rays.n, out.data_ptr(), self._stream()), dev.h,
                "gpis_net_features_1d")
        return out

    def classes(self, rays: RayBatch, reply):
        import torch
        out = torch.empty(rays.n, dtype=torch.int32, device="cuda")
        r = A.ReplySoA()
        for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
            setattr(r, k, reply[k].data_ptr())
        soa = rays.soa()
        self._check(self.L.gpis_net_classes_1d(self.h, C.byref(soa), C.byref(r), rays.n, out.data_ptr(),
                                               self._stream()), "gpis_net_classes_1d")
        return out

    # ---- training ----------------------------------------------------------
    def gradient(self, feats, classes) -> Tuple["object", float]:
        import torch
        g = torch.empty(N_PARAMS, dtype=torch.float32, device="cuda")
        self._check(self.L.gpis_net_gradient(self.h, feats.data_ptr(), classes.data_ptr(), feats.shape[0],
                                             g.data_ptr(), self._loss.data_ptr(), self._stream()),
                    "gpis_net_gradient")
        return g, float(self._loss.item()) / feats.shape[0]

    def train_step(self, feats, classes, lr: float = 1e-3, loss_buf=None):
        """One Adam step; returns the device buffer holding the batch's summed loss."""
        buf = self._loss if loss_buf is None else loss_buf
        self._check(self.L.gpis_net_train_step(self.h, feats.data_ptr(), classes.data_ptr(), feats.shape[0],
                                               lr, buf.data_ptr(), self._stream()), "gpis_net_train_step")
        return buf

    def set_averaging(self, beta: float):
        self._check(self.L.gpis_net_set_averaging(self.h, float(beta)), "gpis_net_set_averaging")

    def use_average(self):
        self._check(self.L.gpis_net_use_average(self.h), "gpis_net_use_average")

    def train(self, feats, classes, epochs: int = 20, batch: int = 65536, lr: float = 1e-3, seed: int = SEED_TRAIN,
              log=None, average: float = 0.0) -> List[float]:
        """Shuffled mini-batch Adam over the whole set; returns the mean loss of each epoch.
        `average` > 0: the network ends with the exponential moving average (that beta) of
        its Adam iterates, which removes the step noise of the constant learning rate."""
        if average > 0.0:
            self.set_averaging(average)
        import torch
        n = (feats.shape[0] // batch) * batch
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        steps = n // batch
        losses = torch.zeros(epochs * steps, dtype=torch.float32, device="cuda")
        out = []
        for ep in range(epochs):
            perm = torch.randperm(feats.shape[0], device="cuda", generator=gen)[:n]
            for i in range(steps):
                idx = perm[i * batch:(i + 1) * batch]
                self.train_step(feats.index_select(0, idx), classes.index_select(0, idx), lr,
                                losses[ep * steps + i:ep * steps + i + 1])
            m = float(losses[ep * steps:(ep + 1) * steps].sum().item()) / n
            out.append(m)
            if log:
                log(f"epoch {ep + 1}/{epochs}: mean cross-entropy {m:.4f}")
        if average > 0.0:
            self.use_average()
        return out

    # ---- inference ---------------------------------------------------------
    def forward(self, feats):
        import torch
        p = torch.empty((feats.shape[0], N_CLASS), dtype=torch.float32, device="cuda")
        self._check(self.L.gpis_net_forward(self.h, feats.data_ptr(), feats.shape[0], p.data_ptr(), self._stream()),
                    "gpis_net_forward")
        return p

    def flights(self, dev: GpisDevice, rays: RayBatch, reply=None, count_fallback: bool = True,
                no_alloc: bool = False):
        """Surrogate flights with the plain method as fallback; returns (reply, n_fallback).
        `no_alloc`: the allocation-free entry point (at most 262,144 rays per call)."""
        if reply is None:
            reply = dev.alloc_reply(rays.n)
        r = A.ReplySoA()
        for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
            setattr(r, k, reply[k].data_ptr())
        soa = rays.soa()
        nfb = C.c_uint32(0)
        fn = self.L.gpis_surrogate_flights_1d_noalloc if no_alloc else self.L.gpis_surrogate_flights_1d
        st = fn(dev.h, self.h, C.byref(soa), rays.n, C.byref(r), C.byref(nfb) if count_fallback else None,
                self._stream())
        if st != A.GPIS_OK:
            msg = self.L.gpis_net_last_error(self.h).decode(errors="replace") or \
                self.L.gpis_last_error(dev.h).decode(errors="replace")
            raise A.GpisError(f"{fn.__name__}: {A.STATUS_NAMES.get(st, st)} {msg}")
        return reply, int(nfb.value)


# ---------------------------------------------------------------- datasets --
def normalised_distance(rays: RayBatch, reply):
    """s = (t - near) / (far - near) clamped to [0, 1]; exits are 1 (torch, device)."""
    near = rays.t["near_t"]
    far = rays.t["far_t"]
    s = ((reply["t"] - near) / (far - near)).clamp(0.0, 1.0)
    oc = reply["outcome"] & 0xFF
    exited = oc == A.EXIT
    return s.masked_fill(exited, 1.0), exited, oc


def flight_dataset(dev: GpisDevice, net: FlightSurrogate, med: SparseConvolutionNoiseMedium, n: int,
                   seed: int = SEED_TRAIN, chunk: int = 1 << 21):
    """n conditioned 1D flights of the renewal workload traced by the plain method:
    returns device tensors (features n x 8, s, exited, outcome)."""
    import torch
    from . import workloads as W
    seqs = np.random.SeedSequence(seed).spawn((n + chunk - 1) // chunk)
    feats, ss, ex, ocs = [], [], [], []
    done = 0
    for sq in seqs:
        m = min(chunk, n - done)
        rays = W.renewal1d_rays(med, n=m, seed=sq, first_id=done)
        rb = RayBatch.from_numpy(rays)
        rep = dev.trace_flights_1d(rb)
        feats.append(net.features(dev, rb))
        s, e, oc = normalised_distance(rb, rep)
        ss.append(s); ex.append(e); ocs.append(oc)
        done += m
    torch.cuda.synchronize()
    return torch.cat(feats), torch.cat(ss), torch.cat(ex), torch.cat(ocs)


def classes_of(net: FlightSurrogate, s, outcome):
    """Training classes of flights with normalised distances s (through
    gpis_net_classes_1d on a unit bound: near 0, far 1)."""
    import torch
    n = s.shape[0]
    z = torch.zeros(n, dtype=torch.float32, device=s.device)
    soa = A.RaySoA()
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "jitter", "noise_seed", "normal_seed", "flags", "prev_value",
              "prev_gx", "prev_gy", "prev_gz"):
        setattr(soa, k, z.data_ptr())
    ones = torch.ones_like(z)
    soa.near_t, soa.far_t, soa.length_scale = z.data_ptr(), ones.data_ptr(), None
    sc = s.contiguous()
    oc = outcome.to(torch.int32).contiguous()
    r = A.ReplySoA()
    r.t, r.outcome = sc.data_ptr(), oc.data_ptr()
    for k in ("value", "gx", "gy", "gz", "evals"):
        setattr(r, k, z.data_ptr())
    out = torch.empty(n, dtype=torch.int32, device=s.device)
    net._check(net.L.gpis_net_classes_1d(net.h, C.byref(soa), C.byref(r), n, out.data_ptr(), net._stream()),
               "gpis_net_classes_1d")
    return out


# ---------------------------------------------------------------- judging --
def ks_two_sample(a: np.ndarray, b: np.ndarray) -> float:
    a = np.sort(np.asarray(a, np.float64))
    b = np.sort(np.asarray(b, np.float64))
    grid = np.concatenate([a, b])
    fa = np.searchsorted(a, grid, side="right") / len(a)
    fb = np.searchsorted(b, grid, side="right") / len(b)
    return float(np.max(np.abs(fa - fb)))


def strata_of(keys: np.ndarray, bins: int = 4) -> np.ndarray:
    """Stratum index of each row: the quantile bin (of `bins`) of every column of keys."""
    keys = np.asarray(keys, np.float64)
    sid = np.zeros(keys.shape[0], np.int64)
    for j in range(keys.shape[1]):
        qs = np.quantile(keys[:, j], np.linspace(0.0, 1.0, bins + 1)[1:-1])
        sid = sid * bins + np.searchsorted(qs, keys[:, j], side="right")
    return sid


def stratified_ks(sid: np.ndarray, a: np.ndarray, b: np.ndarray, min_count: int = 2000) -> Dict[str, float]:
    """Two-sample KS distance between a and b (paired rows) inside each stratum: the
    count-weighted mean over strata, the largest, and the number of strata used."""
    order = np.argsort(sid, kind="stable")
    sids = sid[order]
    cuts = np.flatnonzero(np.diff(sids)) + 1
    tot, wsum, worst, used = 0.0, 0, 0.0, 0
    for grp in np.split(order, cuts):
        if len(grp) < min_count:
            continue
        k = ks_two_sample(a[grp], b[grp])
        tot += k * len(grp)
        wsum += len(grp)
        worst = max(worst, k)
        used += 1
    return {"mean": tot / max(wsum, 1), "max": worst, "strata": used}


def image_bias(samples_a: np.ndarray, samples_b: np.ndarray) -> Tuple[float, float]:
    """Noise-corrected relative RMS bias between two images given as (pixels, spp) samples,
    and the bias below which it cannot tell the images apart.  With per-pixel means m and
    unbiased sample variances v, d2 = mean over pixels of (ma - mb)^2 - va/spp - vb/spp has
    expectation mean((bias)^2) whatever the noise; returns sqrt(max(d2, 0)) / mean(mb) and
    sqrt(2 * standard error of d2) / mean(mb)."""
    a = np.asarray(samples_a, np.float64)
    b = np.asarray(samples_b, np.float64)
    ma, mb = a.mean(1), b.mean(1)
    va, vb = a.var(1, ddof=1), b.var(1, ddof=1)
    terms = (ma - mb) ** 2 - va / a.shape[1] - vb / b.shape[1]
    d2 = float(terms.mean())
    se = float(terms.std() / math.sqrt(len(terms)))
    scale = max(abs(float(mb.mean())), 1e-12)
    return math.sqrt(max(d2, 0.0)) / scale, math.sqrt(2.0 * se) / scale


def pit_ks(probs: np.ndarray, edges: np.ndarray, s: np.ndarray, exited: np.ndarray, rng) -> float:
    """KS distance from uniform of the probability integral transform of plain-method
    distances under the network's CDFs (randomised over the exit mass)."""
    p = np.asarray(probs, np.float64)
    c = np.cumsum(p, axis=1)
    e = np.asarray(edges, np.float64)
    sv = np.clip(np.asarray(s, np.float64), 0.0, np.nextafter(1.0, 0.0))
    b = np.where(sv > 0, np.searchsorted(e[1:N_BINS], sv, side="right"), 0)
    idx = np.arange(len(sv))
    before = np.where(b > 0, c[idx, np.maximum(b - 1, 0)], 0.0)
    width = e[b + 1] - e[b]
    # inside a bin: linear; in the atom (zero width): randomised over its mass
    frac = np.where(width > 0, (sv - e[b]) / np.where(width > 0, width, 1.0), rng.uniform(size=len(sv)))
    u = before + frac * p[idx, b]
    cross = c[:, N_BINS - 1]
    u = np.where(exited, cross + rng.uniform(size=len(sv)) * (1.0 - cross), u)
    u = np.sort(np.clip(u, 0.0, 1.0))
    n = len(u)
    return float(max(np.max(np.arange(1, n + 1) / n - u), np.max(u - np.arange(n) / n)))


def entropy_of(probs):
    import torch
    p = probs.clamp_min(1e-30)
    return -(probs * p.log()).sum(1)
