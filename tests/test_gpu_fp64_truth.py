"""The GPU's FP32 error against the FP64 truth build of the oracle (ORC_FP64_TRUTH): value
and gradient of the stationary field at 65,536 points (C0 workload 1), outcomes and hit
distances of 65,536 flights from a sphere (C0 workload 2) and of 65,536 conditioned 1D
flights with per-ray length scales (C3), with the endpoint gradient.  Each statistic is
printed beside the float oracle's own error against the same truth."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_desc, oracle_field, oracle_rays

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
NTH = min(16, os.cpu_count() or 1)


def _pct(x):
    return f"median {np.median(x):.2e} p99 {np.percentile(x, 99):.2e} p99.9 {np.percentile(x, 99.9):.2e} max {x.max():.2e}"


def _flight_errors(name, got, truth):
    oc = got["outcome"] & 0xFF
    same = oc == truth["outcome"]
    cr = same & (truth["outcome"] != po.EXIT)
    # relative to max(t, 0.01): a crossing cut off to t = 0 on one side counts as its
    # distance in lattice steps
    rel = np.abs(got["t"].astype(np.float64) - truth["t"]) / np.maximum(np.abs(truth["t"]), 0.01)
    print(f"{name}: outcome agreement {same.mean():.6f}; hit distance err / max(t, 0.01) {_pct(rel[cr])}")
    return same, cr, rel


def test_c0_points_and_flights_fp32_error_against_fp64_truth():
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W
    print("case: C0 65,536 points and 65,536 sphere flights, GPU and float oracle against FP64 truth", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(W.config0_json())
    F, T = oracle_field(med), po.Truth64(oracle_desc(med))
    pts = W.points_workload(65536)
    dev = GpisDevice(med)
    x = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    v, gx, gy, gz = dev.eval_points(x[:, 0].contiguous(), x[:, 1].contiguous(), x[:, 2].contiguous(), med.seed)
    rays = W.sphere_rays(med, n=65536)
    rep = dev.trace_flights(RayBatch.from_numpy(rays))
    torch.cuda.synchronize()
    gpu = torch.stack([v, gx, gy, gz], 1).cpu().numpy().astype(np.float64)
    got = {k: t.cpu().numpy() for k, t in rep.items()}
    dev.close()
    truth = T.eval_points(pts, med.seed, nthreads=NTH)
    f32 = F.eval_points(pts, med.seed, nthreads=NTH).astype(np.float64)
    gmag = np.maximum(np.linalg.norm(truth[:, 1:], axis=1), 1.0)
    for name, a in (("GPU", gpu), ("float oracle", f32)):
        dv = np.abs(a[:, 0] - truth[:, 0])
        dg = np.linalg.norm(a[:, 1:] - truth[:, 1:], axis=1) / gmag
        print(f"C0 points, {name} against FP64: value abs err {_pct(dv)}; gradient err / max(|g|, 1) {_pct(dg)}; "
              f"support flips (|dv| > 1e-4) {(dv > 1e-4).mean():.5f}")
        if name == "GPU":
            gdv, gdg = dv, dg
    # FP32 rounding of the splat sums, except where |to|^2 rounds across the support radius
    # (one truncated splat, ~1e-2): at most 1 point in 10^3
    assert np.median(gdv) < 2e-6 and np.median(gdg) < 2e-6
    assert (gdv > 1e-4).mean() < 1e-3 and (gdg > 1e-4).mean() < 1e-3
    orr = oracle_rays(rays, med)
    t64 = T.trace(orr, nthreads=NTH)
    same, cr, rel = _flight_errors("C0 sphere flights, GPU against FP64", got, t64)
    _flight_errors("C0 sphere flights, float oracle against FP64", F.trace(orr, nthreads=NTH), t64)
    assert same.mean() >= 0.999
    assert np.median(rel[cr]) < 1e-6 and np.percentile(rel[cr], 99) < 1e-4


def test_c3_flights_fp32_error_against_fp64_truth():
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W
    print("case: C3 65,536 1D flights, GPU and float oracle against FP64 truth", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    desc = oracle_desc(med)
    rays = W.renewal1d_rays(med, n=65536)
    dev = GpisDevice(med)
    rep = dev.trace_flights_1d(RayBatch.from_numpy(rays))
    torch.cuda.synchronize()
    got = {k: t.cpu().numpy() for k, t in rep.items()}
    dev.close()
    orr = oracle_rays(rays, med)
    ls = rays["length_scale"]
    t64 = po.trace_ls64(desc, orr, ls, nthreads=NTH)
    same, cr, rel = _flight_errors("C3 1D flights, GPU against FP64", got, t64)
    _flight_errors("C3 1D flights, float oracle against FP64", po.trace_ls(desc, orr, ls, nthreads=NTH), t64)
    assert same.mean() >= 0.999
    assert np.median(rel[cr]) < 1e-6 and np.percentile(rel[cr], 99) < 1e-4
    # the endpoint gradient of a hit: the truth's conditioned realization at the GPU's point
    hit = np.nonzero(cr & (t64["outcome"] == po.HIT))[0]
    t = got["t"][hit].astype(np.float64)
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1)[hit].astype(np.float64)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)[hit].astype(np.float64)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    q = (o + t[:, None] * d).astype(np.float32)
    ref = po.conditioned_eval_ls64(desc, orr[hit], ls[hit], q, t)
    g = np.stack([got["gx"], got["gy"], got["gz"]], 1)[hit].astype(np.float64)
    dg = np.linalg.norm(g - ref[:, 1:], axis=1) / np.maximum(np.linalg.norm(ref[:, 1:], axis=1), 1.0)
    print(f"C3 hit gradient, GPU against FP64 at the GPU's point: err / max(|g|, 1) {_pct(dg)} over {hit.size} hits")
    assert np.median(dg) < 1e-5 and np.percentile(dg, 99.9) < 1e-3
