"""The "noise" variance and length-scale fields of a proc_nonstationary covariance
(ProceduralNoise / ProceduralNoiseVec, GPFunctions.hpp:647-830, GPFunctions.cpp:53-139)
on the HIP path against the oracle on the same seeded inputs: field values and gradients
at fixed points for every noise type, in the multi-resolution blend and in the
brute-force local-scale splats, and 3D flights of the medium of the reference's
shaderball-nonstationary scene (data/example-scenes/classic-global/
shaderball-nonstationary.json: knob mean, multi-resolution grid, bottom_top variance and
length-scale fields) through the parity gate, unconditioned and under Renewal+."""
import os

import numpy as np
import pytest

from tests.helpers import oracle_field, oracle_rays
from tests.parity import parity_gate

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
KNOB = {"position": [0.0, 0.8, 0.0], "scale": 1.0, "rotation": [0.0, 0.0, 0.0]}


def _shaderball(multires=True):
    """The shaderball-nonstationary scene's Gaussian process in a sparse_conv_noise medium."""
    from gpis_extend_amd import workloads as W
    js = W.with_procedural_mean(W.config0_json(), "knob", KNOB)
    js["gaussian_process"]["covariance"] = {
        "type": "proc_nonstationary", "multiResolutionGrid": multires,
        "cov": {"type": "squared_exponential", "sigma": 1, "lengthScale": 1},
        "var": {"type": "noise", "noise": "bottom_top", "min": 0.05, "max": 0.0016, "start": 0, "end": 0.6},
        "ls": {"type": "noise", "noise": "bottom_top", "min": 0.1, "max": 0.032, "start": 0, "end": 0.6},
    }
    js.update({"single_realization": False, "correlation_context": "global", "seed": 5, "step_size": 0.005})
    return js


def _fields(var, ls, multires):
    from gpis_extend_amd import workloads as W
    js = W.config2_json()
    cov = js["gaussian_process"]["covariance"]
    cov["multiResolutionGrid"] = multires
    cov["var"] = var
    cov["ls"] = ls
    js["single_realization"] = True
    return js


CASES = [
    ("bottom_top", {"type": "noise", "noise": "bottom_top", "min": 0.05, "max": 0.2, "start": -0.5, "end": 0.6},
     {"type": "noise", "noise": "bottom_top", "min": 0.1, "max": 0.032, "start": 0, "end": 0.6}),
    ("left_right/front_back", {"type": "noise", "noise": "left_right", "min": 0.5, "max": 1.5, "start": -1, "end": 1},
     {"type": "noise", "noise": "front_back", "min": 0.15, "max": 0.04, "start": -0.8, "end": 0.8}),
    ("bottom_top_left_right", {"type": "noise", "noise": "bottom_top_left_right", "min": 0.3, "max": 1.2,
                               "min2": 0.4, "max2": 1.0, "start": -1, "end": 1, "start2": -1, "end2": 1},
     {"type": "noise", "noise": "bottom_top_left_right", "min": 0.06, "max": 0.1, "min2": 0.5, "max2": 1.0,
      "start": -1, "end": 1, "start2": -1, "end2": 1}),
    ("sandstone/rust", {"type": "noise", "noise": "sandstone", "min": 0.3, "max": 1.5},
     {"type": "noise", "noise": "rust"}),
]


@pytest.mark.parametrize("multires", [True, False])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_eval_points_noise_fields(case, multires):
    from gpis_extend_amd import GpisDevice, SparseConvolutionNoiseMedium, workloads as W
    name, var, ls = CASES[case]
    print(f"case: eval_points, noise fields {name}, multiResolutionGrid={multires}, 8192 points", flush=True)
    js = _fields(var, ls, multires)
    med = SparseConvolutionNoiseMedium.from_json(js)
    dev = GpisDevice(med)
    x = W.points_workload(8192, seed=W.SEED_POINTS)
    xt = torch.from_numpy(x).cuda()
    v, gx, gy, gz = dev.eval_points(xt[:, 0].contiguous(), xt[:, 1].contiguous(), xt[:, 2].contiguous(), med.seed)
    got = torch.stack([v, gx, gy, gz], 1).cpu().numpy()
    dev.close()
    ref = oracle_field(med).eval_points(x, med.seed, nthreads=NTH)
    dv = np.abs(got[:, 0] - ref[:, 0])
    gmag = np.linalg.norm(ref[:, 1:], axis=1)
    dg = np.linalg.norm(got[:, 1:] - ref[:, 1:], axis=1) / np.maximum(gmag, 1.0)
    off = (dv > 1e-4) | (dg > 1e-4)
    print(f"{name}: value max|d| {dv.max():.3e}, grad max|d|/max(|g|,1) {dg.max():.3e}, beyond 1e-4: {off.sum()}")
    # the field's variance and scale are the oracle's float values wherever its double
    # arithmetic rounds alike; a point may differ only by the FP32 noise sum
    assert off.mean() <= 2e-4, off.sum()


def _flights(js, rays):
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium
    med = SparseConvolutionNoiseMedium.from_json(js)
    dev = GpisDevice(med)
    rep = dev.trace_flights(RayBatch.from_numpy(rays))
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in rep.items()}
    dev.close()
    F = oracle_field(med)
    orays = oracle_rays(rays, med)
    return got, F.trace(orays, nthreads=NTH), (F, orays)


def test_flights_shaderball_nonstationary_primary():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print("case: 3D flights, shaderball-nonstationary medium, 2048 rays from radius 2.2", flush=True)
    js = _shaderball()
    med = SparseConvolutionNoiseMedium.from_json(js)
    rays = W.sphere_rays(med, n=2048, radius=2.2, seed=31)
    got, ref, orc = _flights(js, rays)
    oc = got["outcome"] & 0xFF
    print(f"hits {(oc == 1).sum()}, exits {(oc == 0).sum()}")
    assert (oc == 1).sum() > 200
    parity_gate(got, ref, "shaderball-nonstationary primary", min_agree=0.999, oracle=orc)


def test_flights_shaderball_nonstationary_renewal_plus():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print("case: 3D flights, shaderball-nonstationary medium under Renewal+, 1024 rays", flush=True)
    js = _shaderball()
    js["correlation_context"] = "renewal+"
    med = SparseConvolutionNoiseMedium.from_json(js)
    rays = W.sphere_rays(med, n=1024, radius=2.2, seed=32)
    rays["flags"] = np.zeros_like(rays["flags"])
    rng = np.random.default_rng(33)
    rays["prev_value"] = rng.normal(0.3, 0.1, len(rays["ox"])).astype(np.float32)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)
    g = (-d * rng.uniform(0.5, 1.5, (len(d), 1))).astype(np.float32)
    rays["prev_gx"], rays["prev_gy"], rays["prev_gz"] = g[:, 0].copy(), g[:, 1].copy(), g[:, 2].copy()
    got, ref, orc = _flights(js, rays)
    parity_gate(got, ref, "shaderball-nonstationary Renewal+", min_agree=0.99, oracle=orc)


def test_noise_field_refusals():
    from gpis_extend_amd import SparseConvolutionNoiseMedium
    from gpis_extend_amd.medium import MediumConfigError
    js = _shaderball()
    js["gaussian_process"]["covariance"]["ls"]["noise"] = "marble"
    with pytest.raises(MediumConfigError, match="Invalid noise typ function"):
        SparseConvolutionNoiseMedium.from_json(js)
