"""Device-resident 1D paths (gpis_trace_paths_1d) on the HIP path against the oracle's
restatement of the same paths (orc_trace_paths1d) on the same seeded inputs: the
hit-time phase (mirror and Lambert BSDFs in the frame of the sampled gradient), the
light sample with its conditioned shadow flight, and Renewal+ chaining from hit to hit.

Every flight of a path is the 1D flight whose parity is pinned elsewhere (bit-identical
distances), so whole paths agree flight for flight: counts, outcomes, the last flight's
distance and the accumulated light samples.  The tolerances cover the FP32 rounding of
values that the oracle forms with a different libm (double log/cos/sin of the Box-Muller
draws, float cos/sin of the Lambert sample)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_desc, oracle_rays

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
BOX = 1.0


def _medium(js):
    from gpis_extend_amd import SparseConvolutionNoiseMedium
    return SparseConvolutionNoiseMedium.from_json(js)


def _gpu_paths(med, rays, max_bounces, phase, light, bounce_seed=7):
    from gpis_extend_amd import GpisDevice, RayBatch
    dev = GpisDevice(med)
    # the reference's rounding: every flight of a path bit-identical with the oracle's
    dev.set_precision("exact")
    rb = RayBatch.from_numpy(rays)
    dev.reset_counters()
    rep = dev.trace_paths_1d(rb, rays["pss"], rays["scene_seed"], max_bounces, phase=phase,
                             bounce_seed=bounce_seed, box=BOX, light=light)
    torch.cuda.synchronize()
    counters = dev.counters()
    fl = dev.trace_flights_1d(rb)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in rep.items() if not k.startswith("_")}
    single = {k: v.cpu().numpy() for k, v in fl.items()}
    dev.close()
    return got, single, counters


def _oracle_paths(med, rays, max_bounces, phase, light, bounce_seed=7):
    desc = oracle_desc(med)
    orays = oracle_rays(rays, med)
    kind = po.PHASE_MIRROR if phase == "mirror" else po.PHASE_LAMBERT
    return po.trace_paths1d(desc, orays, rays.get("length_scale"), max_bounces, kind, bounce_seed, BOX, light,
                            nthreads=NTH)


def _compare(name, got, single, ref, counters, min_same, nee_rel=1e-4):
    n = got["flights"].size
    # the first flight of each path is the single flight of the same request: the kernel
    # derives the seeds from pixelSampleSegment and must land on the same realization
    for k in ("t0", "gx0", "gy0", "gz0"):
        assert np.array_equal(got[k].view(np.uint32), single[k[:-1]].view(np.uint32)), k
    assert np.array_equal(got["outcome0"], single["outcome"].view(np.int32))
    # whole paths against the oracle
    oc_last = got["outcome_last"] & 0xFF
    fr_last = (got["outcome_last"] >> 8) & 0xFF
    same = ((got["flights"] == ref["flights"]) & (got["hits"] == ref["hits"]) &
            (got["shadows"] == ref["shadows"]) & (got["visible"] == ref["visible"]) &
            (oc_last == ref["outcome_last"]) & (fr_last == ref["fail_last"]))
    t_ok = np.abs(got["t_last"].astype(np.float64) - ref["t_last"]) <= 1e-4 * np.maximum(1.0, np.abs(ref["t_last"]))
    nee_ok = np.abs(got["nee"].astype(np.float64) - ref["nee"]) <= nee_rel * np.maximum(np.abs(ref["nee"]), 1e-6)
    full = same & t_ok & nee_ok
    bits = np.mean(same & (got["t_last"].view(np.uint32) == ref["t_last"].view(np.uint32)))
    print(f"{name}: {n} paths, {got['flights'].sum()} flights, {got['shadows'].sum()} shadow flights; "
          f"paths agreeing {full.mean():.5f} (structure {same.mean():.5f}, last t bit-identical {bits:.5f}); "
          f"mean nee gpu {got['nee'].mean():.6g} oracle {ref['nee'].mean():.6g}", flush=True)
    assert full.mean() >= min_same, full.mean()
    # the device counters see every flight the paths traced, shadow flights included
    assert counters["evals"] == int(got["evals"].astype(np.int64).sum())
    assert counters["invalid"] == 0
    return full


@pytest.mark.parametrize("case", ["renewal+", "renewal+ correlationXY", "renewal"])
def test_mirror_paths_with_light_samples(case):
    from gpis_extend_amd import workloads as W
    print(f"case: 1D mirror paths, {case}, 16,384 paths of up to 5 flights + light samples", flush=True)
    js = W.config3_json()
    if case == "renewal":
        js["correlation_context"] = "renewal"
    if case.endswith("correlationXY"):
        js["1D_gradient_correlationXY"] = True
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=16384, seed=21)
    got, single, counters = _gpu_paths(med, rays, 4, "mirror", W.LIGHT_POS)
    ref = _oracle_paths(med, rays, 4, "mirror", W.LIGHT_POS)
    _compare(f"mirror paths ({case})", got, single, ref, counters, min_same=0.999)
    assert got["hits"].sum() > got["flights"].size  # the workload does chain hits
    assert got["shadows"].sum() > 0 and got["visible"].sum() > 0


def test_lambert_paths_with_light_samples():
    from gpis_extend_amd import workloads as W
    print("case: 1D Lambert paths, Renewal+, 16,384 paths of up to 5 flights + light samples", flush=True)
    js = W.config3_json()
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=16384, seed=22)
    got, single, counters = _gpu_paths(med, rays, 4, "lambert", W.LIGHT_POS)
    ref = _oracle_paths(med, rays, 4, "lambert", W.LIGHT_POS)
    # the Lambert direction goes through float cos/sin (device libm against the C
    # library): a last-bit difference moves later flights by rounding, rarely more
    _compare("Lambert paths", got, single, ref, counters, min_same=0.99)


def test_paths_without_light_and_the_medium_bounce_limit():
    from gpis_extend_amd import workloads as W
    print("case: 1D mirror paths, no light, medium max_bounces 3, 8,192 paths", flush=True)
    js = W.config3_json()
    js["max_bounces"] = 3
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=8192, seed=23)
    got, single, counters = _gpu_paths(med, rays, 6, "mirror", None)
    ref = _oracle_paths(med, rays, 6, "mirror", None)
    full = _compare("mirror paths, bounce limit", got, single, ref, counters, min_same=0.999)
    assert np.all(got["shadows"] == 0) and np.all(got["nee"] == 0.0)
    # flight w = 3 of a path fails before its march (beginDistance's bounce limit)
    lim = ((got["outcome_last"] & 0xFF) == po.FAIL) & (((got["outcome_last"] >> 8) & 0xFF) == po.FAIL_BOUNCE_LIMIT)
    assert lim.sum() > 0
    assert np.all(got["flights"][lim] == 4) and np.all(got["hits"][lim] == 3)
    assert np.all(got["flights"] <= 4)
    assert full[lim].all()


def test_paths_refuse_bad_arguments():
    from gpis_extend_amd import GpisDevice, RayBatch, workloads as W
    from gpis_extend_amd._capi import GpisError
    print("case: 1D path entry point refusals", flush=True)
    med1 = _medium(W.config3_json())
    rays = W.renewal1d_rays(med1, n=64, seed=24)
    dev = GpisDevice(med1)
    rb = RayBatch.from_numpy(rays)
    with pytest.raises(GpisError, match="INVALID_ARGUMENT"):
        dev.trace_paths_1d(rb, rays["pss"], 3, 65)
    with pytest.raises(GpisError, match="INVALID_ARGUMENT"):
        dev.trace_paths_1d(rb, rays["pss"], 3, 2, box=0.0)
    dev.close()
    med3 = _medium(W.config0_json())
    r3 = W.sphere_rays(med3, n=64)
    d3 = GpisDevice(med3)
    with pytest.raises(GpisError, match="INVALID_ARGUMENT"):
        d3.trace_paths_1d(RayBatch.from_numpy(r3), r3["pss"], 0, 2)
    d3.close()
