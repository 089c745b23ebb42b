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
