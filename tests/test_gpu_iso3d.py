"""3D flights in the isotropic ray space (isotropic_3D_sampling) on the HIP path against
the oracle on the same seeded inputs: primary flights with an anisotropic SE kernel,
Renewal+ conditioned flights from a surface, paths whose first flight equals the single
flight, and the entry points that refuse the mode."""
import os

import numpy as np
import pytest

from tests.helpers import oracle_field, oracle_rays
from tests.parity import parity_gate

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)


def _js(ctxt="global", single=True, aniso=(1.0, 1.5, 0.8)):
    from gpis_extend_amd import workloads as W
    js = W.config0_json()
    js["isotropic_3D_sampling"] = True
    js["single_realization"] = single
    js["correlation_context"] = ctxt
    js["gaussian_process"]["covariance"]["aniso"] = list(aniso)
    return js


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


def test_iso_ray_primary_flights():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print("case: isotropic ray space, anisotropic SE, 8,192 primary flights", flush=True)
    js = _js()
    rays = W.sphere_rays(SparseConvolutionNoiseMedium.from_json(js), n=8192)
    got, ref, orc = _flights(js, rays)
    parity_gate(got, ref, "isotropic ray space (8,192 flights)", oracle=orc)


def test_iso_ray_renewal_plus_conditioned_flights():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    print("case: isotropic ray space, Renewal+, 2,048 surface-born flights", flush=True)
    js = _js(ctxt="renewal+", single=False)
    rays = W.box_surface_rays(SparseConvolutionNoiseMedium.from_json(js), n=2048)
    got, ref, orc = _flights(js, rays)
    # surface-born flights start at f ~ 0: their first sign decisions are grazing by construction
    parity_gate(got, ref, "isotropic ray space, Renewal+ (2,048 flights)", min_agree=0.99, oracle=orc)


def test_iso_ray_paths_first_flight_and_refusals():
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W
    from gpis_extend_amd._capi import GpisError
    print("case: isotropic ray space paths (2,048 camera rays) and refusals", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(_js())
    dev = GpisDevice(med)
    rays = W.camera_rays(med, width=64, height=64)
    rb = RayBatch.from_numpy(rays)
    fl = dev.trace_flights(rb)
    pa = dev.trace_paths(rb, 2, 7, 1.0)
    torch.cuda.synchronize()
    assert torch.equal(fl["t"], pa["t0"])
    assert torch.equal(fl["outcome"], pa["outcome0"])
    with pytest.raises(GpisError, match="UNSUPPORTED"):
        dev.set_domain((-1.0,) * 3, (1.0,) * 3)
    x = torch.zeros(16, device="cuda")
    with pytest.raises(GpisError, match="UNSUPPORTED"):
        dev.eval_points(x, x, x, med.seed)
    dev.close()
