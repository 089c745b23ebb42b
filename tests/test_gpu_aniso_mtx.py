"""The full anisotropy matrix of the SE covariance ("useAnisoMtx" and "anisoMtx",
GPFunctions.cpp:663-681, 706-764, 784-843) on the HIP path against the oracle on the same
seeded inputs: field values and gradients at fixed points with impulses from the generator
and from the packed table, primary flights and flights conditioned under Renewal+ (the
full-matrix covariance splats and the inverse of its second derivative) through the parity
gate, in world space and in the isotropic ray space."""
import math
import os

import numpy as np
import pytest

from tests.helpers import oracle_field, oracle_rays
from tests.parity import parity_gate

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)


def _rot(ax, ang):
    ax = np.asarray(ax, np.float64)
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


MATS = {
    "rotated": (_rot((1, 2, 0.5), 0.7) @ np.diag([1.6, 0.5, 1.0])).astype(np.float32).reshape(9).tolist(),
    "sheared": [1.0, 0.6, 0.0, 0.0, 0.8, -0.3, 0.2, 0.0, 1.2],
}


def _js(name, **kw):
    from gpis_extend_amd import workloads as W
    js = W.config0_json()
    cov = js["gaussian_process"]["covariance"]
    cov["useAnisoMtx"] = True
    cov["anisoMtx"] = MATS[name]
    js.update(kw)
    return js


@pytest.mark.parametrize("table", [False, True])
@pytest.mark.parametrize("name", sorted(MATS))
def test_eval_points_full_matrix(name, table):
    from gpis_extend_amd import GpisDevice, SparseConvolutionNoiseMedium, workloads as W
    print(f"case: eval_points, anisoMtx {name}, table={table}, 8192 points", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(_js(name))
    dev = GpisDevice(med)
    if table:
        dev.set_domain((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    x = W.points_workload(8192, seed=W.SEED_POINTS)
    xt = torch.from_numpy(x).cuda()
    v, gx, gy, gz = dev.eval_points(xt[:, 0].contiguous(), xt[:, 1].contiguous(), xt[:, 2].contiguous(), med.seed)
    got = torch.stack([v, gx, gy, gz], 1).cpu().numpy()
    dev.close()
    ref = oracle_field(med).eval_points(x, med.seed, nthreads=NTH)
    dv = np.abs(got[:, 0] - ref[:, 0])
    gmag = np.linalg.norm(ref[:, 1:], axis=1)
    dg = np.linalg.norm(got[:, 1:] - ref[:, 1:], axis=1) / np.maximum(gmag, 1.0)
    print(f"{name}: value max|d| {dv.max():.3e}, grad max|d|/max(|g|,1) {dg.max():.3e}")
    # FP32: the value to 1e-4 absolute, the gradient to 1e-4 of max(|g|, 1); the support test
    # is the oracle's (|t|^2 < 1 in grid units), so no point may be off by a truncated splat
    assert dv.max() < 1e-4
    assert dg.max() < 1e-4


def _flights(js, rays, domain=None):
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium
    med = SparseConvolutionNoiseMedium.from_json(js)
    dev = GpisDevice(med)
    if domain is not None:
        dev.set_domain(*domain)
    rep = dev.trace_flights(RayBatch.from_numpy(rays))
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in rep.items()}
    dev.close()
    F = oracle_field(med)
    orays = oracle_rays(rays, med)
    return got, F.trace(orays, nthreads=NTH), (F, orays)


@pytest.mark.parametrize("name", sorted(MATS))
def test_flights_full_matrix_primary(name):
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print(f"case: 3D flights, anisoMtx {name}, 8192 rays from radius 2 with a table", flush=True)
    js = _js(name)
    med = SparseConvolutionNoiseMedium.from_json(js)
    rays = W.sphere_rays(med, n=8192, seed=41)
    got, ref, orc = _flights(js, rays, ((-2.1,) * 3, (2.1,) * 3))
    oc = got["outcome"] & 0xFF
    print(f"hits {(oc == 1).sum()}, exits {(oc == 0).sum()}")
    assert (oc == 1).sum() > 1000
    # camera-like primary flights: a few per ten thousand graze the surface (the gate's
    # grazing test must name every disagreement)
    parity_gate(got, ref, f"anisoMtx {name} primary", min_agree=0.999, oracle=orc)


@pytest.mark.parametrize("name", sorted(MATS))
def test_flights_full_matrix_renewal_plus(name):
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print(f"case: 3D flights, anisoMtx {name} under Renewal+, 2048 rays", flush=True)
    js = _js(name, single_realization=False, correlation_context="renewal+", seed=5)
    med = SparseConvolutionNoiseMedium.from_json(js)
    rays = W.sphere_rays(med, n=2048, radius=2.0, seed=42)
    rays["flags"] = np.zeros_like(rays["flags"])
    rng = np.random.default_rng(43)
    rays["prev_value"] = rng.normal(0.3, 0.1, len(rays["ox"])).astype(np.float32)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)
    g = (-d * rng.uniform(0.5, 1.5, (len(d), 1)) + rng.normal(0, 0.3, d.shape)).astype(np.float32)
    rays["prev_gx"], rays["prev_gy"], rays["prev_gz"] = g[:, 0].copy(), g[:, 1].copy(), g[:, 2].copy()
    got, ref, orc = _flights(js, rays)
    parity_gate(got, ref, f"anisoMtx {name} Renewal+", min_agree=0.99, oracle=orc)



@pytest.mark.parametrize("name", sorted(MATS))
def test_ray_space_flights_full_matrix(name):
    """The full matrix in the isotropic ray space (isotropic_3D_sampling): the frame of
    normalize(W d), points at frame.toLocal(W p), gradients W^T frame.toGlobal g; primary
    flights and Renewal+ flights from a surface (L^T delta into the ray space)."""
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print(f"case: isotropic ray space, anisoMtx {name}: 8,192 primary and 2,048 Renewal+ flights", flush=True)
    js = _js(name, isotropic_3D_sampling=True)
    rays = W.sphere_rays(SparseConvolutionNoiseMedium.from_json(js), n=8192, seed=44)
    got, ref, orc = _flights(js, rays)
    parity_gate(got, ref, f"ray space anisoMtx {name} primary", min_agree=0.999, oracle=orc)
    js = _js(name, isotropic_3D_sampling=True, correlation_context="renewal+", single_realization=False)
    rays = W.box_surface_rays(SparseConvolutionNoiseMedium.from_json(js), n=2048)
    got, ref, orc = _flights(js, rays)
    # surface-born flights start at f ~ 0: their first sign decisions are grazing by construction
    parity_gate(got, ref, f"ray space anisoMtx {name} Renewal+", min_agree=0.99, oracle=orc)
