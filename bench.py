"""Flights/s of the sparse-convolution GPIS flight path on one MI355X.

Workload (one "step"): the primary-ray-flights-at-scale configuration -- a
stationary SE field (sigma 1, lengthScale 0.05, 8 Bernoulli impulses per cell,
single realization, seed 1) around a sphere SDF of radius 0.5; 2048x2048
pinhole camera rays (40 deg FOV, eye (0,0,3) looking at the origin) clipped to
the box [-1,1]^3, each traced as a path of one primary flight plus up to 2
cosine-weighted diffuse bounces (bounce seed 7), all in one persistent launch.
Inputs are generated on the host, copied to HBM before timing, and stay there.

The same line carries, under "plain_1d", "paths_1d" and "surrogate", the 1D
Renewal+ flights with NEE (2,097,152 rays), the device-resident 1D mirror paths
with light samples started from the same rays, and the learned surrogate against
the plain method (trained on 2^24 plain flights, served on a 512x512 image at 16
spp), each with its own timing.

Prints one JSON line: the metric and its value, timing, the workload, the roofline of the
dominant kernel and the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Algorithmic FP32 work per 3D value+gradient evaluation, N impulses per cell:
# 27N*8 (pair test) + (4pi/3)N*24 (in-support kernel, gradient, accumulate) + 16.
def flops_per_eval_3d(n_imp: int) -> float:
    return 27 * n_imp * 8 + (4.0 * np.pi / 3.0) * n_imp * 24 + 16


# ... and per 1D evaluation along a ray: 3 cells x N pair tests (4 flops), 2N in-support
# splats (10 flops: kernel, derivative, accumulate), 32 for setup, conditioning, mean.
def flops_per_eval_1d(n_imp: int) -> float:
    return 3 * n_imp * 4 + 2 * n_imp * 10 + 32


PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak (packed FMA issue)
PEAK_FP32_SCALAR_TFLOPS = 78.6    # one FMA per lane per cycle, no packed issue


def cpu_baseline(med, rays, max_seconds=12.0):
    """The oracle (a CPU port of the reference algorithm) on a bounded sample of the
    same workload: whole paths (primary flight + up to 2 bounces) of the first
    camera rays, all host cores, flights/s."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_field, oracle_rays
    from gpis_extend_amd import workloads as W
    from gpis_extend_amd.hashing import Pcg32, xxhash32_4

    F = oracle_field(med)
    cores = min(os.cpu_count() or 1, 64)
    n_total = rays["ox"].shape[0]
    order = np.random.default_rng(12345).permutation(n_total)  # uniform sample over the image
    sample = 2048
    flights = 0
    spent = 0.0
    done_rays = 0
    start_idx = 0
    while spent < max_seconds and start_idx < n_total:
        idx = np.sort(order[start_idx:start_idx + sample])
        start_idx += len(idx)
        cur = W.rays_subset(rays, idx)
        ray_id = idx.astype(np.uint32)
        for bounce in range(3):
            orr = oracle_rays(cur, med)
            t0 = time.perf_counter()
            res = F.trace(orr, nthreads=cores)
            spent += time.perf_counter() - t0
            flights += len(res)
            hit = res["outcome"] == po.HIT
            if bounce == 2 or not hit.any():
                break
            # cosine-weighted bounce about the normalised gradient (same rule as the GPU path kernel)
            sel = np.nonzero(hit)[0]
            g = res["grad"][sel].astype(np.float32)
            nrm = g / np.linalg.norm(g, axis=1, keepdims=True)
            h = xxhash32_4(ray_id[sel], np.full(len(sel), bounce + 1, np.uint32), np.full(len(sel), 7, np.uint32),
                           np.zeros(len(sel), np.uint32)) + np.uint32(1)
            pcg = Pcg32(h.astype(np.uint64))
            u1, u2, jit = pcg.next_1d(), pcg.next_1d(), pcg.next_1d()
            r_, ph = np.sqrt(u1), np.float32(2 * np.pi) * u2
            loc = np.stack([r_ * np.cos(ph), r_ * np.sin(ph), np.sqrt(np.maximum(0, 1 - u1))], 1)
            sgn = np.where(nrm[:, 2] < 0, -1.0, 1.0)
            a = -1.0 / (sgn + nrm[:, 2]); b = nrm[:, 0] * nrm[:, 1] * a
            tt = np.stack([1 + sgn * nrm[:, 0] ** 2 * a, sgn * b, -sgn * nrm[:, 0]], 1)
            bb = np.stack([b, sgn + nrm[:, 1] ** 2 * a, -nrm[:, 1]], 1)
            d = loc[:, :1] * tt + loc[:, 1:2] * bb + loc[:, 2:] * nrm
            d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
            o = (np.stack([cur["ox"][sel], cur["oy"][sel], cur["oz"][sel]], 1)
                 + res["t"][sel, None].astype(np.float32) * np.stack([cur["dx"][sel], cur["dy"][sel], cur["dz"][sel]], 1)).astype(np.float32)
            _, tf = W.box_interval(o.astype(np.float64), d.astype(np.float64), 1.0)
            nxt = W._empty_rays(len(sel))
            nxt["ox"], nxt["oy"], nxt["oz"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
            nxt["dx"], nxt["dy"], nxt["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
            nxt["far_t"] = tf.astype(np.float32)
            nxt["jitter"] = jit.astype(np.float32)
            nxt["flags"] = np.zeros(len(sel), np.uint32)
            nxt["prev_gx"], nxt["prev_gy"], nxt["prev_gz"] = g[:, 0].copy(), g[:, 1].copy(), g[:, 2].copy()
            nxt["pss"] = cur["pss"][sel]
            nxt["noise_seed"] = cur["noise_seed"][sel]
            nxt["normal_seed"] = cur["normal_seed"][sel]
            cur = nxt
            ray_id = ray_id[sel]
        done_rays += len(idx)
    return {"value": flights / spent, "unit": "flights/s", "cores": cores, "kind": "port",
            "sample": f"{done_rays} camera-ray paths (primary + up to 2 bounces, {flights} flights) of the same "
                      f"workload through the C oracle on {cores} host threads, {spent:.1f} s"}


def cpu_baseline_1d(med, rays, light, max_seconds=12.0):
    """The oracle on a bounded uniform sample of the 1D workload: conditioned 1D flights
    with their per-ray length scales, then NEE toward the light at every hit."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_desc, oracle_rays
    from gpis_extend_amd import workloads as W

    desc = oracle_desc(med)
    cores = min(os.cpu_count() or 1, 64)
    n_total = rays["ox"].shape[0]
    order = np.random.default_rng(12345).permutation(n_total)
    sample, flights, spent, start_idx = 8192, 0, 0.0, 0
    while spent < max_seconds and start_idx < n_total:
        idx = np.sort(order[start_idx:start_idx + sample])
        start_idx += len(idx)
        cur = W.rays_subset(rays, idx)
        orr = oracle_rays(cur, med)
        t0 = time.perf_counter()
        res = po.trace_ls(desc, orr, cur["length_scale"], nthreads=cores)
        hit = np.nonzero(res["outcome"] == po.HIT)[0]
        t = res["t"][hit].astype(np.float32)
        p = np.stack([cur["ox"][hit] + t * cur["dx"][hit], cur["oy"][hit] + t * cur["dy"][hit],
                      cur["oz"][hit] + t * cur["dz"][hit]], 1).astype(np.float32)
        wl = np.asarray(light, np.float32)[None, :] - p
        wl = wl / np.linalg.norm(wl, axis=1, keepdims=True)
        d = np.stack([cur["dx"][hit], cur["dy"][hit], cur["dz"][hit]], 1)
        h = (wl - d) / np.linalg.norm(wl - d, axis=1, keepdims=True)
        po.nee_ls(desc, orr[hit], cur["length_scale"][hit], h.astype(np.float32), p, t)
        spent += time.perf_counter() - t0
        flights += len(idx)
    return {"value": flights / spent, "unit": "flights/s", "cores": cores, "kind": "port",
            "sample": f"{flights} conditioned 1D flights (+ NEE at each hit) of the same workload through "
                      f"the C oracle, flights on {cores} host threads and NEE on 1, {spent:.1f} s"}


def run_1d(args):
    """1D along-ray Renewal+ flights with NEE toward a point light (one step = one
    launch over all rays)."""
    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W

    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    rays = W.renewal1d_rays(med, n=args.rays_1d)
    n = rays["ox"].shape[0]
    dev = GpisDevice(med)
    batch = RayBatch.from_numpy(rays)
    reply = dev.alloc_reply(n)
    stream = torch.cuda.current_stream()
    if args.waves_per_cu:
        dev.set_launch(args.waves_per_cu, 0)
    for _ in range(args.warmup):
        dev.trace_flights_1d(batch, light=W.LIGHT_POS, reply=reply)
    torch.cuda.synchronize()
    dev.reset_counters()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        dev.trace_flights_1d(batch, light=W.LIGHT_POS, reply=reply)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))
    c = dev.counters()
    flights_per_step = c["flights"] / args.steps
    evals_per_step = c["evals"] / args.steps
    ms_per_step = wall * 1e3 / args.steps
    fpe = flops_per_eval_1d(int(med.impulse_density))
    achieved = evals_per_step * fpe / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "GPIS 1D flights/s on 1 GPU (plain method, NEE at each hit)",
        "value": flights_per_step / (ms_per_step / 1e3), "unit": "flights/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain 1D along-ray Renewal+ flights: SE sigma 1, l ~ U[0.02, 0.2] per ray, "
                               "sphere SDF mean, v0 ~ N(0.3, 0.2^2), directional derivative ~ N(0, 1), "
                               "NEE toward a point light at (0, 2, 0)",
                   "rays": n, "flights_per_step": flights_per_step, "evals_per_step": evals_per_step,
                   "hits_per_step": c["hits"] / args.steps, "fails_per_step": c["fails"] / args.steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "frac_scalar_issue_peak": achieved / PEAK_FP32_SCALAR_TFLOPS,
                     "kernel": "flight1d_kernel<8>", "kernel_ms": kernel_ms, "flops_per_eval": fpe},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_1d(med, rays, W.LIGHT_POS)
    dev.close()
    return line


def cpu_baseline_paths_1d(med, rays, light, max_bounces, max_seconds=12.0):
    """The oracle's path restatement on a bounded uniform sample of the same paths."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_desc, oracle_rays
    from gpis_extend_amd import workloads as W

    desc = oracle_desc(med)
    cores = min(os.cpu_count() or 1, 64)
    n_total = rays["ox"].shape[0]
    order = np.random.default_rng(12345).permutation(n_total)
    sample, paths, spent, start_idx = 8192, 0, 0.0, 0
    while spent < max_seconds and start_idx < n_total:
        idx = np.sort(order[start_idx:start_idx + sample])
        start_idx += len(idx)
        cur = W.rays_subset(rays, idx)
        orr = oracle_rays(cur, med)
        t0 = time.perf_counter()
        po.trace_paths1d(desc, orr, cur["length_scale"], max_bounces, po.PHASE_MIRROR, 7, 1.0, light,
                         nthreads=cores)
        spent += time.perf_counter() - t0
        paths += len(idx)
    return {"value": paths / spent, "unit": "paths/s", "cores": cores, "kind": "port",
            "sample": f"{paths} mirror paths of the same workload through the C oracle's path restatement "
                      f"on {cores} host threads, {spent:.1f} s"}


def run_paths_1d(args):
    """Device-resident 1D paths on the renewal workload: each ray starts a mirror path of
    up to 5 flights; every hit takes a light sample (NEE normal, pdf and a conditioned
    shadow flight) and bounces, all inside one launch (one step = one launch)."""
    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W

    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    rays = W.renewal1d_rays(med, n=args.rays_1d)
    n = rays["ox"].shape[0]
    dev = GpisDevice(med)
    batch = RayBatch.from_numpy(rays)
    stream = torch.cuda.current_stream()
    max_bounces = 4

    seeds = dev.path_seeds(rays["pss"])
    reply = dev.alloc_path1d_reply(n)

    def step():
        return dev.trace_paths_1d(batch, seeds, rays["scene_seed"], max_bounces, phase="mirror",
                                  bounce_seed=7, box=1.0, light=W.LIGHT_POS, reply=reply)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    dev.reset_counters()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = None
    for i in range(args.steps):
        starts[i].record(stream)
        rep = step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))
    c = dev.counters()
    flights = int(rep["flights"].sum().item())
    shadows = int(rep["shadows"].sum().item())
    evals_per_step = c["evals"] / args.steps
    ms_per_step = wall * 1e3 / args.steps
    fpe = flops_per_eval_1d(int(med.impulse_density))
    achieved = evals_per_step * fpe / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "GPIS 1D paths/s on 1 GPU (mirror paths of up to 5 flights, light sample at each hit)",
        "value": n / (ms_per_step / 1e3), "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1D Renewal+ mirror paths from the renewal workload's rays: up to 1 + 4 flights, "
                               "each hit with neeGrad/neePDF toward (0, 2, 0) and a conditioned shadow flight, "
                               "bounces inside [-1, 1]^3",
                   "paths": n, "flights_per_step": flights, "shadow_flights_per_step": shadows,
                   "evals_per_step": evals_per_step,
                   "flights_per_s": (flights + shadows) / (ms_per_step / 1e3)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "flight1d_kernel<8,true>", "kernel_ms": kernel_ms, "flops_per_eval": fpe},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_paths_1d(med, rays, W.LIGHT_POS, max_bounces)
    dev.close()
    return line


MLP_FLOPS_PER_ROW = 2 * (8 * 128 + 3 * 128 * 128 + 128 * 32)  # 108,544 (forward)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def run_surrogate(args):
    """The learned flight sampler: train on plain-method flights of the renewal
    workload (seed 11), then serve fresh rays arranged as a 512x512 image with 16
    samples per pixel (seed 12), against the plain method on the same rays."""
    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W
    from gpis_extend_amd import surrogate as S

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    dev = GpisDevice(med)
    net = S.FlightSurrogate(seed=S.SEED_TRAIN)
    t0 = time.perf_counter()
    feats, s, ex, oc = S.flight_dataset(dev, net, med, args.train_flights, seed=S.SEED_TRAIN)
    log(f"training set: {feats.shape[0]} plain flights in {time.perf_counter() - t0:.1f} s")
    cal = S.calibrate(feats, s[~ex])
    net.set_calibration(cal)
    cls = S.classes_of(net, s, oc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    losses = net.train(feats, cls, epochs=args.epochs, batch=65536, lr=1e-3, seed=S.SEED_TRAIN, log=log)
    e1.record()
    torch.cuda.synchronize()
    train_s = e0.elapsed_time(e1) / 1e3
    samples = (feats.shape[0] // 65536) * 65536 * args.epochs
    # entropy bound: the 99.5th percentile of the trained network's entropy on training inputs
    sub = feats[: min(feats.shape[0], 1 << 20)]
    H = S.entropy_of(net.forward(sub))
    cal.entropy_max = float(torch.quantile(H.float(), 0.995).item())
    net.set_calibration(cal)
    del feats, s, ex, oc, cls
    torch.cuda.empty_cache()

    rays = W.renewal1d_image_rays(med, width=512, spp=16, seed=S.SEED_INFER)
    rb = RayBatch.from_numpy(rays)
    n = rb.n
    plain = dev.alloc_reply(n)
    sur = dev.alloc_reply(n)
    for _ in range(args.warmup):
        dev.trace_flights_1d(rb, reply=plain)
        net.flights(dev, rb, reply=sur, count_fallback=False)
    torch.cuda.synchronize()

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        a.record()
        for _ in range(args.steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return (time.perf_counter() - w0) / args.steps, a.elapsed_time(b) / args.steps

    plain_wall, plain_ms = timed(lambda: dev.trace_flights_1d(rb, reply=plain))
    sur_wall, sur_ms = timed(lambda: net.flights(dev, rb, reply=sur, count_fallback=False))
    _, nfb = net.flights(dev, rb, reply=sur)
    fe = net.features(dev, rb)
    _, fwd_ms = timed(lambda: net.forward(fe))
    torch.cuda.synchronize()

    sp, exp_, _ = S.normalised_distance(rb, plain)
    ss, _, _ = S.normalised_distance(rb, sur)
    sp_np, ss_np = sp.cpu().numpy(), ss.cpu().numpy()
    ks = S.ks_two_sample(sp_np, ss_np)
    # PIT of the plain distances under the network's CDFs, on the rays it serves
    served = (sur["evals"] == 0).cpu().numpy()
    probs = net.forward(fe).cpu().numpy()
    rng = np.random.default_rng(5)
    pit = S.pit_ks(probs[served], cal.edges, sp_np[served], exp_.cpu().numpy()[served], rng)
    # image error: per-pixel mean normalised distance over the 16 samples
    img_p = sp_np.reshape(-1, 16).mean(1)
    img_s = ss_np.reshape(-1, 16).mean(1)
    rays2 = W.renewal1d_image_rays(med, width=512, spp=16, seed=S.SEED_INFER, scene_seed=4)
    rb2 = RayBatch.from_numpy(rays2)
    plain2 = dev.trace_flights_1d(rb2)
    sp2, _, _ = S.normalised_distance(rb2, plain2)
    img_p2 = sp2.cpu().numpy().reshape(-1, 16).mean(1)
    rel = lambda a, b: float(np.mean(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)))
    line = {
        "metric": "GPIS 1D flights/s on 1 GPU, surrogate (plain fallback) vs plain",
        "value": n / sur_wall, "unit": "flights/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sur_wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "renewal 1D flights (SE sigma 1, l ~ U[0.02, 0.2] per ray, sphere SDF mean, "
                               "v0 ~ N(0.3, 0.2^2), derivative ~ N(0, 1)) as a 512x512 image at 16 spp; "
                               "MLP 8-128x4-32 trained on plain flights",
                   "rays": n, "train_flights": int(args.train_flights), "epochs": args.epochs, "batch": 65536},
        "plain": {"value": n / plain_wall, "ms_per_step": plain_wall * 1e3, "kernel_ms": plain_ms},
        "speedup_vs_plain": plain_wall / sur_wall,
        "fallback_fraction": nfb / n,
        "ks_vs_plain": ks, "pit_ks": pit,
        "image_rel_error": rel(img_s, img_p), "image_rel_error_plain_vs_plain": rel(img_p2, img_p),
        "train": {"seconds": train_s, "samples_per_s": samples / train_s, "loss_first_epoch": losses[0],
                  "loss_last_epoch": losses[-1],
                  "tflops": samples * 3 * MLP_FLOPS_PER_ROW / train_s / 1e12},
        "roofline": {"bound": "mfma", "achieved": n * MLP_FLOPS_PER_ROW / (fwd_ms / 1e3) / 1e12,
                     "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": n * MLP_FLOPS_PER_ROW / (fwd_ms / 1e3) / 1e12 / PEAK_BF16_TFLOPS,
                     "traffic": None, "kernel": "infer_kernel (probabilities only)", "kernel_ms": fwd_ms,
                     "flops_per_row": MLP_FLOPS_PER_ROW},
    }
    net.close()
    dev.close()
    return line


def run_two_spheres(args):
    """Primary flights through the sparse_conv_noise medium of the reference's two-spheres
    scene (data/example-scenes/two-spheres-backlight/scene.json): procedural two_spheres
    mean (transform scale 0.1, scale 0.1), SE sigma 0.08, l 0.2, 10 impulses per cell,
    step 0.0025, a fresh realization per flight (no impulse table)."""
    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W

    js = W.with_procedural_mean(W.config0_json(), "two_spheres", {"scale": 0.1}, scale=0.1)
    js["gaussian_process"]["covariance"] = {"type": "squared_exponential", "sigma": 0.08, "lengthScale": 0.2}
    js.update({"step_size": 0.0025, "min_step": 8, "seed": 17429, "impulse_density": 10,
               "single_realization": False, "correlation_context": "global"})
    med = SparseConvolutionNoiseMedium.from_json(js)
    width = min(args.width, 1024)
    rays = W.camera_rays(med, width=width, height=width, eye=(0.0, 0.0, 4.0), box=2.0)
    n = rays["ox"].shape[0]
    dev = GpisDevice(med)
    batch = RayBatch.from_numpy(rays)
    reply = dev.alloc_reply(n)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        dev.trace_flights(batch, reply)
    torch.cuda.synchronize()
    dev.reset_counters()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        dev.trace_flights(batch, reply)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    c = dev.counters()
    dev.close()
    flights = c["flights"] / args.steps
    evals = c["evals"] / args.steps
    ms = wall * 1e3 / args.steps
    return {"metric": "GPIS flights/s on 1 GPU (two-spheres scene medium, procedural SDF mean)",
            "value": flights / (ms / 1e3), "unit": "flights/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 (mean in f64)", "data": "synthetic",
            "config": {"workload": f"{width}x{width} pinhole primary flights from (0, 0, 4) through the "
                                   "two-spheres-backlight medium", "rays": n, "flights_per_step": flights,
                       "evals_per_step": evals, "hits_per_step": c["hits"] / args.steps},
            "kernel_ms": kernel_ms, "evals_per_s": evals / (kernel_ms / 1e3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--workload", choices=("c1", "c3", "c3paths", "c4", "twospheres"), default="c1",
                    help="c1: camera-ray paths on the 3D field (default); c3: 1D Renewal+ flights with NEE; "
                         "c3paths: device-resident 1D mirror paths with light samples; "
                         "c4: the learned surrogate against the plain method; "
                         "twospheres: primary flights through the two-spheres scene medium")
    ap.add_argument("--train-flights", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true",
                    help="report only the camera-ray paths workload (skip the 1D and surrogate legs)")
    ap.add_argument("--rays-1d", type=int, default=1 << 21)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape sweep that runs before the warmup steps")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB device-to-device once "
                         "(a known-size kernel for calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(flights shard by ray index: one process per GPU over disjoint ray ranges)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c3":
        print(json.dumps(run_1d(args)), flush=True)
        return
    if args.workload == "c3paths":
        print(json.dumps(run_paths_1d(args)), flush=True)
        return
    if args.workload == "c4":
        print(json.dumps(run_surrogate(args)), flush=True)
        return
    if args.workload == "twospheres":
        print(json.dumps(run_two_spheres(args)), flush=True)
        return

    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W

    med = SparseConvolutionNoiseMedium.from_json(W.config1_json())
    rays = W.camera_rays(med, width=args.width, height=args.width)
    n = rays["ox"].shape[0]
    dev = GpisDevice(med)
    # the impulse table covers every point a flight evaluates: the [-1, 1]^3 box and the
    # camera position (a flight whose secant cuts off at t = 0 takes its gradient at the
    # ray origin)
    dev.set_domain((-1.0, -1.0, -1.0), (1.0, 1.0, 3.0))
    batch = RayBatch.from_numpy(rays)
    reply = dev.alloc_path_reply(n)
    stream = torch.cuda.current_stream()
    tuned = None
    if args.waves_per_cu:
        dev.set_launch(args.waves_per_cu, 0)
    elif not args.no_autotune:
        from gpis_extend_amd import autotune
        tuned = autotune.tune(dev, "paths", n, lambda: dev.trace_paths(batch, 2, 7, 1.0, reply), lists=True)

    for _ in range(args.warmup):
        dev.trace_paths(batch, 2, 7, 1.0, reply)
    torch.cuda.synchronize()
    dev.reset_counters()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        dev.trace_paths(batch, 2, 7, 1.0, reply)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    c = dev.counters()
    flights_per_step = c["flights"] / args.steps
    evals_per_step = c["evals"] / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = flights_per_step / (ms_per_step / 1e3)
    fpe = flops_per_eval_3d(int(med.impulse_density))
    achieved = evals_per_step * fpe / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "GPIS flights/s on 1 GPU (plain method)",
        "value": value,
        "unit": "flights/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain primary-ray flights at scale: SE l=0.05, 8 impulses/cell, "
                               f"{args.width}x{args.width} pinhole rays + 2 diffuse bounces",
                   "rays": n, "flights_per_step": flights_per_step, "evals_per_step": evals_per_step,
                   "hits_per_step": c["hits"] / args.steps, "fails_per_step": c["fails"] / args.steps,
                   "launch": {"refill_batch": tuned[0], "waves_per_cu": tuned[1], "block": tuned[2],
                              "list_subcells": tuned[3], "points_per_list_pass": 2} if tuned else None,
                   "list_passes_per_step": c["passes"] / args.steps, "paired_per_step": c["paired"] / args.steps,
                   "spec_evals_per_step": c["spec_evals"] / args.steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "frac_scalar_issue_peak": achieved / PEAK_FP32_SCALAR_TFLOPS,
                     "kernel": "flight_kernel<8,true>", "kernel_ms": kernel_ms,
                     "flops_per_eval": fpe},
    }
    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(med, rays)
    dev.close()
    if not args.no_extras:
        # the other two headline comparisons: plain 1D flights with NEE, and the learned
        # surrogate against the plain method (its own timing, KS and image error)
        keep = ("value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline", "plain", "speedup_vs_plain",
                "fallback_fraction", "ks_vs_plain", "pit_ks", "image_rel_error", "image_rel_error_plain_vs_plain",
                "train", "metric", "dtype")
        for key, fn in (("plain_1d", run_1d), ("paths_1d", run_paths_1d), ("surrogate", run_surrogate)):
            sub = fn(args)
            line[key] = {k: v for k, v in sub.items() if k in keep}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
