"""The allocation-free surrogate entry point (gpis_surrogate_flights_1d_noalloc) against the
allocating one (gpis_surrogate_flights_1d): the same replies bit for bit on two calls in a
row, the second with fewer rays (the shared scratch keeps the first call's contents), and a
clear refusal of more rays than its fixed scratch holds."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _calibration():
    from gpis_extend_amd.surrogate import Calibration
    big = np.full(8, 3.0e38, np.float32)
    # a finite entropy bound, so that some served rays fall back as well as the forced ones
    return Calibration(np.zeros(8, np.float32), np.ones(8, np.float32), -big, big,
                       np.linspace(0.0, 1.0, 32).astype(np.float32), entropy_max=3.3)


def _rays(med, n, seed):
    from gpis_extend_amd import workloads as W
    rays = W.renewal1d_rays(med, n=n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    # a fifth unconditioned: those fall back to the plain method
    rays["flags"] = np.where(rng.uniform(size=n) < 0.2, 1, 0).astype(np.uint32)
    return rays


def test_noalloc_surrogate_flights_equal_the_allocating_entry_point():
    from gpis_extend_amd import GpisDevice, RayBatch, SparseConvolutionNoiseMedium, workloads as W
    from gpis_extend_amd import surrogate as S
    from gpis_extend_amd._capi import GpisError
    print("case: surrogate flights, allocating vs allocation-free entry point, 200,000 then 3,000 rays", flush=True)
    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    dev = GpisDevice(med)
    net = S.FlightSurrogate(seed=5)
    net.set_calibration(_calibration())
    for n, seed in ((200_000, 91), (3_000, 92)):
        rb = RayBatch.from_numpy(_rays(med, n, seed))
        ref, nfb_ref = net.flights(dev, rb)
        got, nfb = net.flights(dev, rb, no_alloc=True)
        torch.cuda.synchronize()
        print(f"n={n}: fallback {nfb_ref} (allocating) {nfb} (allocation-free)", flush=True)
        assert nfb == nfb_ref
        assert 0 < nfb < n
        for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
            a = ref[k].cpu().numpy()
            b = got[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, k)
    big = RayBatch.from_numpy(_rays(med, (1 << 18) + 1, 93))
    with pytest.raises(GpisError, match="at most 262144 rays"):
        net.flights(dev, big, no_alloc=True)
    net.close()
    dev.close()
