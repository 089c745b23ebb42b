"""The C++ host batch backend (gpis_extend_amd/host, GpRayBatchBackend-shaped) over
the C ABI: replies in request order across several in-flight batches awaited in
reverse order, an invalid request rejected, onComplete once per batch, and results
identical to the direct C-ABI call."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpis_extend_amd import RayBatch, GpisDevice, SparseConvolutionNoiseMedium, workloads as W  # noqa: E402
from gpis_extend_amd import build as B  # noqa: E402


def test_backend_matches_direct_calls(tmp_path):
    med = SparseConvolutionNoiseMedium.from_json(W.config0_json())
    rays = W.sphere_rays(med, n=9000)
    desc = med.to_desc()
    (tmp_path / "desc.bin").write_bytes(bytes(desc))
    n = rays["ox"].shape[0]
    raw = np.zeros((n, 16), np.float32)
    for a, k in enumerate(("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "jitter")):
        raw[:, a] = rays[k]
    raw[:, 9] = rays["noise_seed"].view(np.float32)
    raw[:, 10] = rays["normal_seed"].view(np.float32)
    raw[:, 11] = rays["flags"].view(np.float32)
    raw[:, 12] = rays["prev_value"]
    raw[:, 13], raw[:, 14], raw[:, 15] = rays["prev_gx"], rays["prev_gy"], rays["prev_gz"]
    raw.tofile(tmp_path / "req.bin")
    out = subprocess.run([B.SELFTEST, str(tmp_path / "desc.bin"), str(tmp_path / "req.bin"), str(n),
                          str(tmp_path / "rep.bin"), "-2.1", "2.1", str(med.seed)],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    rep = np.fromfile(tmp_path / "rep.bin", np.float32).reshape(n, 8)
    assert np.all(rep[:, 7] == 1.0)
    dev = GpisDevice(med)
    dev.set_domain((-2.1,) * 3, (2.1,) * 3)
    d = dev.trace_flights(RayBatch.from_numpy(rays))
    torch.cuda.synchronize()
    for col, k in enumerate(("t", "value", "gx", "gy", "gz")):
        assert np.array_equal(rep[:, col], d[k].cpu().numpy()), k
    assert np.array_equal(rep[:, 5].view(np.uint32), d["outcome"].cpu().numpy().view(np.uint32))
    dev.close()


def _write_paths(path, rays):
    n = rays["ox"].shape[0]
    raw = np.zeros((n, 20), np.uint32)
    f = raw.view(np.float32)
    for a, k in enumerate(("ox", "oy", "oz", "dx", "dy", "dz", "near_t", "far_t", "jitter")):
        f[:, a] = rays[k]
    raw[:, 9] = rays["flags"]
    f[:, 10] = rays["prev_value"]
    f[:, 11], f[:, 12], f[:, 13] = rays["prev_gx"], rays["prev_gy"], rays["prev_gz"]
    raw[:, 14:18] = rays["pss"]
    raw.tofile(path)


def _run_pool(tmp_path, med, rays, phase, light, max_bounces=4, pool=3000, batch=1024, precision="fast"):
    (tmp_path / "desc.bin").write_bytes(bytes(med.to_desc()))
    _write_paths(tmp_path / "paths.bin", rays)
    n = rays["ox"].shape[0]
    args = [B.POOLTEST, str(tmp_path / "desc.bin"), str(tmp_path / "paths.bin"), str(n), str(tmp_path / "res.bin"),
            str(max_bounces), str(phase), "7", "1.0", str(int(rays["scene_seed"])), str(pool), str(batch)]
    if light is not None:
        args += [repr(float(x)) for x in light]
    args.append(precision)
    out = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(out.stdout.strip(), out.stderr.strip())
    assert out.returncode == 0
    res = np.fromfile(tmp_path / "res.bin", np.uint32).reshape(n, 12)
    return {"t0": res[:, 0].view(np.float32), "g0": res[:, 1:4].view(np.float32), "outcome0": res[:, 4],
            "flights": res[:, 5], "hits": res[:, 6], "shadows": res[:, 7], "visible": res[:, 8],
            "nee": res[:, 9].view(np.float32), "t_last": res[:, 10].view(np.float32), "outcome_last": res[:, 11]}


def test_path_pool_over_the_backend_equals_device_resident_paths(tmp_path):
    """The wavefront path pool (gpis_extend_amd/host/path_pool.cpp: paths parked on a
    segment request, submitted in batches through submitBatch, resumed from their
    replies, refilled from the queue) traces the same mirror paths as the device-resident
    kernel, bit for bit: the resumption contract loses nothing between segments."""
    print("case: path pool (3,000 in flight, batches of 1,024) vs gpis_trace_paths_1d, 8,192 mirror paths",
          flush=True)
    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    rays = W.renewal1d_rays(med, n=8192, seed=41)
    del rays["length_scale"]  # segment requests carry the medium's own length scale
    pool = _run_pool(tmp_path, med, rays, 0, None)
    dev = GpisDevice(med)
    rep = dev.trace_paths_1d(RayBatch.from_numpy(rays), rays["pss"], rays["scene_seed"], 4, phase="mirror",
                             bounce_seed=7, box=1.0)
    torch.cuda.synchronize()
    d = {k: v.cpu().numpy() for k, v in rep.items() if not k.startswith("_")}
    dev.close()
    for k in ("t0", "t_last", "nee"):
        assert np.array_equal(pool[k].view(np.uint32), d[k].view(np.uint32)), k
    g0 = np.stack([d["gx0"], d["gy0"], d["gz0"]], 1)
    assert np.array_equal(pool["g0"].view(np.uint32), g0.view(np.uint32))
    for k in ("outcome0", "flights", "hits", "shadows", "visible", "outcome_last"):
        assert np.array_equal(pool[k], d[k].view(np.uint32)), k
    assert d["hits"].sum() > d["flights"].size // 2


def test_path_pool_lambert_light_samples_equal_the_oracle(tmp_path):
    """Lambert paths with light samples (Shadow segments conditioned on (0, gradient))
    through the pool equal the oracle's path restatement bit for bit (both draw the
    Lambert direction with the C library's cos/sin)."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_desc, oracle_rays
    print("case: path pool Lambert + light, 8,192 paths, against the oracle", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    rays = W.renewal1d_rays(med, n=8192, seed=42)
    del rays["length_scale"]
    pool = _run_pool(tmp_path, med, rays, 1, W.LIGHT_POS, precision="exact")
    ref = po.trace_paths1d(oracle_desc(med), oracle_rays(rays, med), None, 4, po.PHASE_LAMBERT, 7, 1.0, W.LIGHT_POS,
                           nthreads=min(16, os.cpu_count() or 1))
    same = ((pool["flights"] == ref["flights"]) & (pool["hits"] == ref["hits"]) &
            (pool["shadows"] == ref["shadows"]) & (pool["visible"] == ref["visible"]) &
            ((pool["outcome_last"] & 0xFF) == ref["outcome_last"]) &
            (pool["t_last"].view(np.uint32) == ref["t_last"].view(np.uint32)) &
            (pool["nee"].view(np.uint32) == ref["nee"].view(np.uint32)))
    print(f"pool vs oracle: {same.mean():.5f} of paths identical; shadows {pool['shadows'].sum()}, "
          f"visible {pool['visible'].sum()}", flush=True)
    assert pool["shadows"].sum() > 1000 and pool["visible"].sum() > 100
    # the transverse draws use the device's double log/cos/sin: a last-bit difference there
    # is the only way the two can part
    assert same.mean() >= 0.999
