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


# HBM bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE and WRITE_SIZE runs of
# this bench (KiB per dispatch, scaled by what the counters read for a 256 MiB device copy in
# the same run: FETCH x2.0, WRITE x1.0).  Counters cannot be read from inside the timed process,
# so the figures are the profiled ones, named with the profile summary they come from.
MEASURED_TRAFFIC = {
    "flight_kernel<8,true>": {"bytes": (9.584e6 * 2.0 + 7.369e5) * 1024, "fetch_bytes": 9.584e6 * 2.0 * 1024,
                              "write_bytes": 7.369e5 * 1024, "source": "profiles/flight_kernel_c1_paths_2048.md"},
    "flight1d_kernel<8,fast>": {"bytes": (1.646e5 * 2.0 + 4.630e5) * 1024, "fetch_bytes": 1.646e5 * 2.0 * 1024,
                                "write_bytes": 4.630e5 * 1024, "source": "profiles/flight1d_kernel_c3_2m.md"},
}

PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak (packed FMA issue)
PEAK_FP32_SCALAR_TFLOPS = 78.6    # one FMA per lane per cycle, no packed issue


_ORACLE_BUILD = None


def _oracle_build() -> str:
    """Builds the oracle for this host (-O3 -march=native, the same strict rounding as the
    checking build) into a fresh temporary folder and loads it for the CPU baselines;
    falls back to the portable -O2 build if the compiler is unavailable."""
    global _ORACLE_BUILD
    if _ORACLE_BUILD is None:
        import subprocess
        import tempfile
        from oracle import pyoracle as po
        out = tempfile.mkdtemp(prefix="gpis_oracle_native_")
        try:
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_DIR={out}"],
                           check=True, timeout=180, capture_output=True)
            po.use_library(os.path.join(out, "libgpis_oracle.so"))
            _ORACLE_BUILD = "gcc -O3 -march=native -ffp-contract=off (built on this host)"
        except Exception as e:  # noqa: BLE001 - any build failure keeps the portable build
            _ORACLE_BUILD = f"gcc -O2 -ffp-contract=off (portable build; native build failed: {type(e).__name__})"
    return _ORACLE_BUILD


def _host_cpus(max_threads=16):
    """(threads used, physical cores visible, logical CPUs visible, CPU model): one worker
    per physical core of this process's CPU set, at most `max_threads` (a GPU's share of
    the host)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", c))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(len(cores), max_threads)), len(cores), len(cpus), model


def cpu_rates(prepare, run, n_max, unit, what, target_s=0.8, runs=5, n0=256):
    """The CPU baseline: the oracle on uniform samples of the workload, single-thread and
    with one worker per physical core, each the median of `runs` timed runs after a
    warm-up run on the same sample.  prepare(n) builds the inputs of an n-unit sample
    (untimed); run(inputs, threads) runs it and returns the units done.  Each sample is
    sized from a first run so that one timed run takes about `target_s`."""
    build = _oracle_build()
    threads, phys, logical, model = _host_cpus()
    res = {}
    for label, nth in (("single", 1), ("all", threads)):
        n = min(n_max, n0)
        inp = prepare(n)
        t0 = time.perf_counter()
        run(inp, nth)
        el = time.perf_counter() - t0
        n = int(min(n_max, max(n, n * target_s / max(el, 1e-4))))
        inp = prepare(n)
        run(inp, nth)  # warm-up on the timed sample
        rates = []
        for _ in range(runs):
            t0 = time.perf_counter()
            units = run(inp, nth)
            rates.append(units / (time.perf_counter() - t0))
        res[label] = (float(np.median(rates)), n, units)
    return {"value": res["all"][0], "unit": unit, "cores": threads, "kind": "port",
            "single_thread": res["single"][0], "runs": runs, "statistic": "median",
            "cpu_model": model, "physical_cores_visible": phys, "logical_cpus_visible": logical, "build": build,
            "sample": f"{what}: {res['single'][1]} (single thread, {res['single'][2]} {unit.split('/')[0]}) and "
                      f"{res['all'][1]} ({threads} threads, {res['all'][2]} {unit.split('/')[0]}) drawn uniformly "
                      f"from the workload, through the C oracle; median of {runs} runs after a warm-up run"}


def cpu_baseline(med, rays):
    """The oracle's path restatement (orc_trace_paths3d: the GPU paths' chaining and bounce
    rule) on uniform samples of the camera-ray paths: primary flight + up to 2 bounces."""
    from tests.helpers import oracle_field, oracle_rays
    from gpis_extend_amd import workloads as W

    F = oracle_field(med)
    order = np.random.default_rng(12345).permutation(rays["ox"].shape[0])

    def prepare(n):
        return oracle_rays(W.rays_subset(rays, np.sort(order[:n])), med)

    def run(orr, nth):
        return int(F.trace_paths3d(orr, 2, 7, 1.0, nthreads=nth)["flights"].sum())

    return cpu_rates(prepare, run, len(order), "flights/s", "camera-ray paths (primary + up to 2 bounces)")


def cpu_baseline_1d(med, rays, light):
    """The oracle on uniform samples of the 1D workload: conditioned 1D flights with their
    per-ray length scales, then NEE toward the light at every hit."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_desc, oracle_rays
    from gpis_extend_amd import workloads as W

    desc = oracle_desc(med)
    order = np.random.default_rng(12345).permutation(rays["ox"].shape[0])

    def prepare(n):
        cur = W.rays_subset(rays, np.sort(order[:n]))
        return cur, oracle_rays(cur, med)

    def run(inp, nth):
        cur, orr = inp
        res = po.trace_ls(desc, orr, cur["length_scale"], nthreads=nth)
        hit = np.nonzero(res["outcome"] == po.HIT)[0]
        t = res["t"][hit].astype(np.float32)
        p = np.stack([cur["ox"][hit] + t * cur["dx"][hit], cur["oy"][hit] + t * cur["dy"][hit],
                      cur["oz"][hit] + t * cur["dz"][hit]], 1).astype(np.float32)
        wl = np.asarray(light, np.float32)[None, :] - p
        wl = wl / np.linalg.norm(wl, axis=1, keepdims=True)
        d = np.stack([cur["dx"][hit], cur["dy"][hit], cur["dz"][hit]], 1)
        h = (wl - d) / np.linalg.norm(wl - d, axis=1, keepdims=True)
        po.nee_ls(desc, orr[hit], cur["length_scale"][hit], h.astype(np.float32), p, t)
        return len(orr)

    return cpu_rates(prepare, run, len(order), "flights/s",
                     "conditioned 1D flights with NEE at each hit (flights threaded, NEE on one thread)", n0=2048)


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
    launch = None
    if args.launch:
        v = [int(x) for x in args.launch.split(",")]
        dev.set_refill(v[0])
        dev.set_launch(v[1], 256)
        launch = {"refill_batch": v[0], "waves_per_cu": v[1], "block": 256}
    elif args.waves_per_cu:
        dev.set_launch(args.waves_per_cu, 0)
    elif not args.no_autotune:
        from gpis_extend_amd import autotune
        r, w, _, _ = autotune.tune(dev, "flights_1d", n,
                                   lambda: dev.trace_flights_1d(batch, light=W.LIGHT_POS, reply=reply),
                                   blocks=(256,))
        launch = {"refill_batch": r, "waves_per_cu": w, "block": 256}
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
    # the reference-rounding mode on the same rays, for comparison (not the headline)
    dev.set_precision("exact")
    dev.trace_flights_1d(batch, light=W.LIGHT_POS, reply=reply)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        dev.trace_flights_1d(batch, light=W.LIGHT_POS, reply=reply)
    e1.record(stream)
    torch.cuda.synchronize()
    exact_ms = e0.elapsed_time(e1) / 3
    dev.set_precision("fast")
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
                   "hits_per_step": c["hits"] / args.steps, "fails_per_step": c["fails"] / args.steps,
                   "precision": "fast (FP32 splats and mean)", "launch": launch,
                   "exact_rounding_mode": {"kernel_ms": exact_ms,
                                           "flights_per_s": flights_per_step / (exact_ms / 1e3)}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_TRAFFIC["flight1d_kernel<8,fast>"]["bytes"],
                     "traffic_detail": MEASURED_TRAFFIC["flight1d_kernel<8,fast>"],
                     "hbm_bytes_algorithmic": n * 112,
                     "frac_scalar_issue_peak": achieved / PEAK_FP32_SCALAR_TFLOPS,
                     "kernel": "flight1d_kernel<8,fast>", "kernel_ms": kernel_ms, "flops_per_eval": fpe},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_1d(med, rays, W.LIGHT_POS)
    dev.close()
    return line


def cpu_baseline_paths_1d(med, rays, light, max_bounces):
    """The oracle's path restatement on uniform samples of the same 1D paths."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_desc, oracle_rays
    from gpis_extend_amd import workloads as W

    desc = oracle_desc(med)
    order = np.random.default_rng(12345).permutation(rays["ox"].shape[0])

    def prepare(n):
        cur = W.rays_subset(rays, np.sort(order[:n]))
        return cur, oracle_rays(cur, med)

    def run(inp, nth):
        cur, orr = inp
        po.trace_paths1d(desc, orr, cur["length_scale"], max_bounces, po.PHASE_MIRROR, 7, 1.0, light, nthreads=nth)
        return len(orr)

    return cpu_rates(prepare, run, len(order), "paths/s", "1D mirror paths with light samples", n0=1024)


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


def _timed_steps(args, step, stream):
    """Warm-up steps, then `args.steps` timed steps: (wall ms per step, mean HIP-event ms
    per step on `stream`)."""
    import torch
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall * 1e3 / args.steps, float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))


def run_c2(args):
    """The nonstationary Renewal+ chains: 1,048,576 rays born on the box surface, each a
    device-resident chain of up to 16 flights (gpis_trace_paths), every flight conditioned
    on (0, the previous hit's gradient), per-flight seeds Vec4u(rayId, 0, 0, flight)."""
    import torch
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W

    med = SparseConvolutionNoiseMedium.from_json(W.config2_json())
    rays = W.box_surface_rays(med, n=args.rays_c2)
    n = rays["ox"].shape[0]
    dev = GpisDevice(med)
    batch = RayBatch.from_numpy(rays)
    seeds = dev.path_seeds(rays["pss"])
    reply = dev.alloc_path_reply(n)
    stream = torch.cuda.current_stream()
    max_bounces = 15

    def step():
        dev.trace_paths(batch, max_bounces, 7, 1.0, reply, pss=seeds, scene_seed=int(rays["scene_seed"]))

    tuned = None
    if not args.no_autotune:
        from gpis_extend_amd import autotune
        tuned = autotune.tune(dev, "chains", n, step)
    dev.reset_counters()
    ms_per_step, kernel_ms = _timed_steps(args, step, stream)
    c = dev.counters()
    steps = args.steps + args.warmup  # the counters run from the reset, warm-up included
    flights_per_step = c["flights"] / steps
    evals_per_step = c["evals"] / steps
    fpe = 2 * flops_per_eval_3d(int(med.impulse_density)) + 40 + 60
    achieved = evals_per_step * fpe / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "GPIS flights/s on 1 GPU (nonstationary Renewal+ chains)",
        "value": flights_per_step / (ms_per_step / 1e3), "unit": "flights/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "nonstationary multi-resolution Renewal+ chains: box SDF (half extent 0.6), "
                               "procedural l(x), origins uniform on the box surface, up to 16 conditioned flights "
                               "per ray, diffuse bounces inside [-1, 1]^3",
                   "rays": n, "flights_per_step": flights_per_step, "evals_per_step": evals_per_step,
                   "hits_per_step": c["hits"] / steps, "fails_per_step": c["fails"] / steps,
                   "launch": {"refill_batch": tuned[0], "waves_per_cu": tuned[1], "block": tuned[2]} if tuned else None},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": f"flight_kernel<{int(med.impulse_density)},true,COND,PATHS>", "kernel_ms": kernel_ms,
                     "flops_per_eval": fpe},
    }
    if not args.no_cpu_baseline:
        from tests.helpers import oracle_field, oracle_rays
        F = oracle_field(med)
        order = np.random.default_rng(12345).permutation(n)

        def prepare(k):
            return oracle_rays(W.rays_subset(rays, np.sort(order[:k])), med)

        def run(orr, nth):
            return int(F.trace_paths3d(orr, max_bounces, 7, 1.0, seeded=True, nthreads=nth)["flights"].sum())

        line["cpu_baseline"] = cpu_rates(prepare, run, n, "flights/s", "Renewal+ chains of up to 16 flights")
    dev.close()
    return line


def run_points(args):
    """Value + gradient of the stationary field (C0: SE l = 0.1, 8 impulses per cell,
    single realization) at 2^20 points uniform in [-1, 1]^3: one gpis_eval_points launch."""
    import torch
    from gpis_extend_amd import GpisDevice, SparseConvolutionNoiseMedium, workloads as W

    med = SparseConvolutionNoiseMedium.from_json(W.config0_json())
    pts = W.points_workload(args.points)
    n = pts.shape[0]
    dev = GpisDevice(med)
    x = torch.from_numpy(np.ascontiguousarray(pts[:, 0])).cuda()
    y = torch.from_numpy(np.ascontiguousarray(pts[:, 1])).cuda()
    z = torch.from_numpy(np.ascontiguousarray(pts[:, 2])).cuda()
    stream = torch.cuda.current_stream()

    # the impulse table over the points' box, and the support lists' sub-cell count (0: no
    # lists) that evaluates fastest
    dev.set_domain((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))

    def step():
        dev.eval_points(x, y, z, med.seed)

    best = None
    for sub in ((0, 3, 4, 6, 8) if not args.no_autotune else (3,)):
        dev.set_lists(sub, 2)
        step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        if best is None or ms < best[1]:
            best = (sub, ms)
    dev.set_lists(best[0], 2)
    ms_per_step, kernel_ms = _timed_steps(args, step, stream)
    fpe = flops_per_eval_3d(int(med.impulse_density))
    achieved = n * fpe / (kernel_ms / 1e3) / 1e12
    line = {
        "metric": "GPIS field evaluations/s on 1 GPU (value + gradient)",
        "value": n / (ms_per_step / 1e3), "unit": "evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "stationary SE field (sigma 1, l 0.1, 8 impulses per cell, single realization, "
                               "seed 1): value and gradient at points uniform in [-1, 1]^3",
                   "points": n, "list_subcells": best[0]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "hbm_bytes_algorithmic": n * 28, "kernel": "eval_points_kernel<8>", "kernel_ms": kernel_ms,
                     "flops_per_eval": fpe},
    }
    if not args.no_cpu_baseline:
        from tests.helpers import oracle_field
        F = oracle_field(med)

        def prepare(k):
            return np.ascontiguousarray(pts[:k])

        def run(xyz, nth):
            F.eval_points(xyz, med.seed, nthreads=nth)
            return xyz.shape[0]

        line["cpu_baseline"] = cpu_rates(prepare, run, n, "evaluations/s", "points of the same workload",
                                         n0=4096)
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
    cal = S.calibrate(feats, s, ex)
    net.set_calibration(cal)
    cls = S.classes_of(net, s, oc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    losses = net.train(feats, cls, epochs=args.epochs, batch=65536, lr=1e-3, seed=S.SEED_TRAIN, log=log,
                       average=S.AVERAGE_BETA)
    e1.record()
    torch.cuda.synchronize()
    train_s = e0.elapsed_time(e1) / 1e3
    samples = (feats.shape[0] // 65536) * 65536 * args.epochs
    # entropy bound: the 99.9th percentile of the trained network's entropy on training inputs
    sub = feats[: min(feats.shape[0], 1 << 20)]
    H = S.entropy_of(net.forward(sub))
    cal.entropy_max = float(torch.quantile(H.float(), S.ENTROPY_QUANTILE).item())
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

    sp, exp_, ocp = S.normalised_distance(rb, plain)
    ss, _, ocs = S.normalised_distance(rb, sur)
    sp_np, ss_np = sp.cpu().numpy(), ss.cpu().numpy()
    ks = S.ks_two_sample(sp_np, ss_np)
    # PIT of the plain distances under the network's CDFs, on the rays it serves
    served = (sur["evals"] == 0).cpu().numpy()
    probs_t = net.forward(fe)
    probs = probs_t.cpu().numpy()
    rng = np.random.default_rng(5)
    pit = S.pit_ks(probs[served], cal.edges, sp_np[served], exp_.cpu().numpy()[served], rng)
    # a second plain render of the same rays with independent realizations: the noise floor
    rays2 = W.renewal1d_image_rays(med, width=512, spp=16, seed=S.SEED_INFER, scene_seed=4)
    rb2 = RayBatch.from_numpy(rays2)
    plain2 = dev.trace_flights_1d(rb2)
    sp2, _, ocp2 = S.normalised_distance(rb2, plain2)
    sp2_np = sp2.cpu().numpy()
    # the conditional test: KS inside the 27 quantile strata of (value, derivative, length
    # scale), against the plain-vs-plain floor and the input-free sampler (the pooled law)
    F = fe.cpu().numpy()
    sid = S.strata_of(F[:, list(S.EDGE_STRATA_FEATURES)], 3)
    ks_strata = S.stratified_ks(sid, sp_np, ss_np)
    ks_floor = S.stratified_ks(sid, sp_np, sp2_np)
    ks_free = S.stratified_ks(sid, sp_np, sp_np[np.random.default_rng(6).permutation(n)])
    # outcome classes of the same rays, and the cross-entropy of the plain classes under
    # the network against the input-free predictor's (the marginal class frequencies)
    frac = lambda o: {k: float((o == v).float().mean().item()) for k, v in (("hit", 1), ("exit", 0), ("fail", 2))}
    cls = S.classes_of(net, sp, ocp)
    ok = cls >= 0
    ce_net = float(-(probs_t[ok].gather(1, cls[ok].long()[:, None]).clamp_min(1e-30).log()).mean().item())
    pm = torch.bincount(cls[ok].long(), minlength=S.N_CLASS).double()
    pm = pm / pm.sum()
    ce_free = float(-(pm[pm > 0] * pm[pm > 0].log()).sum().item())
    # image error: per-pixel mean normalised distance over the 16 samples
    img_p = sp_np.reshape(-1, 16).mean(1)
    img_s = ss_np.reshape(-1, 16).mean(1)
    img_p2 = sp2_np.reshape(-1, 16).mean(1)
    rel = lambda a, b: float(np.mean(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)))
    # the bias test at many samples per pixel: a 128x128 image at 1024 spp (seed 13), the
    # surrogate and a second plain render against the plain render, noise-corrected
    del plain2, rb2, sur
    torch.cuda.empty_cache()
    hi = {}
    for key, scene in (("plain", 3), ("plain2", 4), ("surrogate", 3)):
        rh = RayBatch.from_numpy(W.renewal1d_image_rays(med, width=128, spp=1024, seed=13, scene_seed=scene))
        rep = net.flights(dev, rh)[0] if key == "surrogate" else dev.trace_flights_1d(rh)
        hi[key] = S.normalised_distance(rh, rep)[0].cpu().numpy().reshape(-1, 1024)
        del rh, rep
        torch.cuda.empty_cache()
    bias_s, limit_s = S.image_bias(hi["surrogate"], hi["plain"])
    bias_p, limit_p = S.image_bias(hi["plain2"], hi["plain"])
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
        "ks_strata": ks_strata, "ks_strata_plain_vs_plain": ks_floor, "ks_strata_input_free": ks_free,
        "outcomes": {"plain": frac(ocp), "surrogate": frac(ocs), "plain_second_render": frac(ocp2)},
        "cross_entropy": {"network": ce_net, "input_free": ce_free},
        "image_rel_error": rel(img_s, img_p), "image_rel_error_plain_vs_plain": rel(img_p2, img_p),
        "image_rel_bias": {"image": "128x128 at 1024 spp", "surrogate_vs_plain": bias_s,
                           "plain_vs_plain": bias_p, "detection_limit": max(limit_s, limit_p)},
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


def _copy_probe(args):
    """One device-to-device copy of a known size (a reference kernel for calibrating the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters of the same run)."""
    if not args.copy_probe_mib:
        return
    import torch
    src = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--workload", choices=("c1", "c2", "c0w1", "c3", "c3paths", "c4", "twospheres"), default="c1",
                    help="c1: camera-ray paths on the 3D field (default); c2: nonstationary Renewal+ chains; "
                         "c0w1: field value + gradient at 2^20 points; c3: 1D Renewal+ flights with NEE; "
                         "c3paths: device-resident 1D mirror paths with light samples; "
                         "c4: the learned surrogate against the plain method; "
                         "twospheres: primary flights through the two-spheres scene medium")
    ap.add_argument("--train-flights", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true",
                    help="report only the camera-ray paths workload (skip the 1D and surrogate legs)")
    ap.add_argument("--rays-1d", type=int, default=1 << 21)
    ap.add_argument("--rays-c2", type=int, default=1 << 20)
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape sweep that runs before the warmup steps")
    ap.add_argument("--launch", default="",
                    help="R,W,B[,S]: refill batch, waves per CU, threads per block and support-list "
                         "sub-cells for the camera-ray paths (R,W for the 1D flights, whose blocks are "
                         "256 threads); replaces the autotuner's sweep")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB device-to-device once "
                         "(a known-size kernel for calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(flights shard by ray index: one process per GPU over disjoint ray ranges)", file=sys.stderr)
        sys.exit(2)

    if not args.no_cpu_baseline:
        _oracle_build()  # before anything loads the oracle: the baselines time the host build
    if args.workload == "c3":
        print(json.dumps(run_1d(args)), flush=True)
        _copy_probe(args)
        return
    if args.workload == "c2":
        print(json.dumps(run_c2(args)), flush=True)
        _copy_probe(args)
        return
    if args.workload == "c0w1":
        print(json.dumps(run_points(args)), flush=True)
        _copy_probe(args)
        return
    if args.workload == "c3paths":
        print(json.dumps(run_paths_1d(args)), flush=True)
        _copy_probe(args)
        return
    if args.workload == "c4":
        print(json.dumps(run_surrogate(args)), flush=True)
        _copy_probe(args)
        return
    if args.workload == "twospheres":
        print(json.dumps(run_two_spheres(args)), flush=True)
        _copy_probe(args)
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
    if args.launch:
        v = [int(x) for x in args.launch.split(",")]
        tuned = (v[0], v[1], v[2], v[3] if len(v) > 3 else 3)
        from gpis_extend_amd import autotune
        autotune._apply(dev, tuned, True)
    elif args.waves_per_cu:
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
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_TRAFFIC["flight_kernel<8,true>"]["bytes"],
                     "traffic_detail": MEASURED_TRAFFIC["flight_kernel<8,true>"],
                     "hbm_bytes_algorithmic": n * 96,
                     "frac_scalar_issue_peak": achieved / PEAK_FP32_SCALAR_TFLOPS,
                     "kernel": "flight_kernel<8,true>", "kernel_ms": kernel_ms,
                     "flops_per_eval": fpe},
    }
    _copy_probe(args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(med, rays)
    dev.close()
    if not args.no_extras:
        # the other configurations, each with its own timing, roofline and CPU baseline:
        # the nonstationary chains, the field at 2^20 points, plain 1D flights with NEE,
        # 1D paths, and the learned surrogate against the plain method (KS, image error)
        keep = ("value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline", "plain", "speedup_vs_plain",
                "fallback_fraction", "ks_vs_plain", "pit_ks", "ks_strata", "ks_strata_plain_vs_plain",
                "ks_strata_input_free", "outcomes", "cross_entropy", "image_rel_error",
                "image_rel_error_plain_vs_plain", "image_rel_bias", "train", "metric", "dtype")
        for key, fn in (("chains_c2", run_c2), ("points_c0w1", run_points), ("plain_1d", run_1d),
                        ("paths_1d", run_paths_1d), ("surrogate", run_surrogate)):
            sub = fn(args)
            line[key] = {k: v for k, v in sub.items() if k in keep}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
