"""1D along-ray flights (1D_sampling media) and their NEE on the HIP path against
the oracle on the same seeded inputs: flight outcomes and hit distances through
the parity gate, hit gradients and exit values against the oracle's conditioned
realization at the same point, NEE gradient and pdf toward the point light, and
the per-flight evaluation counts."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_desc, oracle_rays
from tests.parity import EPS_F, parity_gate

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)


def _medium(js, grid=None):
    from gpis_extend_amd import SparseConvolutionNoiseMedium
    return SparseConvolutionNoiseMedium.from_json(js, mean_grid=grid)


def _run(js, rays, light, grid=None, precision="exact"):
    from gpis_extend_amd import GpisDevice, RayBatch
    med = _medium(js, grid)
    dev = GpisDevice(med)
    dev.set_precision(precision)
    rep = dev.trace_flights_1d(RayBatch.from_numpy(rays), light=light)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in rep.items()}
    counters = dev.counters()
    dev.close()
    return med, got, counters


def _fragile_fn(desc, orays, ls, ref, rel_tol=1e-4, trials=16):
    def f(i):
        di = type(desc).from_buffer_copy(desc)
        if ls is not None:
            di.length_scale = float(ls[i])
        F = po.Field(di)
        for s in range(trials):
            for eps in (EPS_F, -EPS_F):
                r = F.trace_perturbed(orays[i], eps, 1000 + s)
                if r["outcome"] != ref[i]["outcome"]:
                    return True
                if abs(r["t"] - ref[i]["t"]) > rel_tol * abs(ref[i]["t"]) + 1e-6:
                    return True
        return False
    return f


def _grad_point(rays, t):
    """ro + normalize(dir in double) * t, rounded to float (the gradient position), with
    the normalisation as a multiplication by 1/|d|: the same point to the last bit, which
    a tabulated mean's forward-difference gradient needs (its float rounding noise is
    ~1e-4 of |grad| between neighbouring points)."""
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1).astype(np.float64)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    l2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    d = d * (1.0 / np.sqrt(l2))[:, None]
    return (o + d * t.astype(np.float64)[:, None]).astype(np.float32)


def _scatter_point(rays, t):
    """origin + t * dir in float (the medium sample point of a hit)."""
    t = t.astype(np.float32)
    return np.stack([rays["ox"] + t * rays["dx"], rays["oy"] + t * rays["dy"], rays["oz"] + t * rays["dz"]],
                    1).astype(np.float32)


def _normalize_f32(v):
    l2 = v[:, 0] * v[:, 0]
    l2 = l2 + v[:, 1] * v[:, 1]
    l2 = l2 + v[:, 2] * v[:, 2]
    inv = np.float32(1.0) / np.sqrt(l2)
    return v * inv[:, None]


def _mirror_normal(rays, p, light):
    L = np.asarray(light, np.float32)[None, :]
    wl = _normalize_f32((L - p).astype(np.float32))
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float32)
    return _normalize_f32((wl - d).astype(np.float32))


def _check_1d(js, rays, light, name, min_agree=0.9999, grid=None, precision="exact"):
    """precision "exact": the reference's rounding (bit-identical distances for most
    flights, identical evaluation counts); "fast": the FP32 path, through the same gate."""
    med, got, counters = _run(js, rays, light, grid, precision)
    desc = oracle_desc(med)
    orays = oracle_rays(rays, med)
    ls = rays.get("length_scale")
    ref = po.trace_ls(desc, orays, ls, nthreads=NTH)
    bad = parity_gate(got, ref, name, min_agree=min_agree, oracle=_fragile_fn(desc, orays, ls, ref))
    oc = got["outcome"] & 0xFF
    agree = (oc == ref["outcome"]) & ~bad
    n = oc.size
    # the 1D evaluation follows the reference's rounding step by step: report how many
    # flights land on the oracle's distance to the last bit
    same_bits = np.mean(got["t"].view(np.uint32) == ref["t"].astype(np.float32).view(np.uint32))
    print(f"{name}: bit-identical t for {same_bits:.6f} of flights")

    # per-flight evaluations: the oracle's count, plus one conditioning evaluation
    # for a conditioned flight and one NEE evaluation per hit when a light is given
    cond = (rays["flags"] & 1) == 0
    extra = cond.astype(np.int64) + ((oc == po.HIT) & (light is not None))
    same_evals = got["evals"][agree].astype(np.int64) == ref["evals"][agree] + extra[agree]
    if precision == "exact":
        assert same_evals.all()
    else:
        # a secant candidate within FP32 rounding of f = 0 may take one more shrink step
        assert same_evals.mean() >= 0.999
    assert counters["flights"] == n
    assert counters["evals"] == int(got["evals"].astype(np.int64).sum())
    assert counters["hits"] == int((oc == po.HIT).sum())

    # hit gradients and exit gradients: the oracle's conditioned realization at the
    # point the HIP path evaluated (same t), t_segment = t
    sel = np.nonzero(agree & ((oc == po.HIT) | (oc == po.EXIT)))[0]
    # the point from the oracle's double t: the HIP path forms it from its own double t,
    # which the float t in the reply only rounds
    q = _grad_point(rays, ref["t"])[sel]
    sub = orays[sel]
    ref4 = po.conditioned_eval_ls(desc, sub, None if ls is None else ls[sel], q, got["t"][sel].astype(np.float32))
    g_gpu = np.stack([got["gx"], got["gy"], got["gz"]], 1)[sel].astype(np.float64)
    g_ref = ref4[:, 1:].astype(np.float64)
    gmag = np.linalg.norm(g_ref, axis=1)
    gerr = np.linalg.norm(g_gpu - g_ref, axis=1) / np.maximum(gmag, 1.0)
    print(f"{name}: gradient max rel err {gerr.max():.3e} (median {np.median(gerr):.1e}) over {sel.size} flights")
    # FP32 tolerance for the gradient: 1e-4 relative to max(|g|, 1)
    assert gerr.max() < 1e-4

    if light is None:
        assert "nee_pdf" not in got
        return got, ref
    hits = np.nonzero(agree & (oc == po.HIT))[0]
    p = _scatter_point(rays, got["t"])[hits]
    h = _mirror_normal({k: rays[k][hits] for k in ("dx", "dy", "dz")}, p, light)
    gr, pdf = po.nee_ls(desc, orays[hits], None if ls is None else ls[hits], h, p, got["t"][hits].astype(np.float32))
    ng = np.stack([got["nee_gx"], got["nee_gy"], got["nee_gz"]], 1)[hits].astype(np.float64)
    nerr = np.linalg.norm(ng - gr, axis=1) / np.maximum(np.linalg.norm(gr, axis=1), 1.0)
    gp = got["nee_pdf"][hits].astype(np.float64)
    rp = pdf.astype(np.float64)
    perr = np.abs(gp - rp) / np.maximum(np.abs(rp), 1e-30)
    print(f"{name}: NEE over {hits.size} hits: grad max rel err {nerr.max():.3e}; pdf max rel err "
          f"{perr[rp > 1e-6].max() if (rp > 1e-6).any() else 0:.3e}; pdf>0 for {(rp > 0).mean():.3f}")
    # NEE in FP32: the gradient to 1e-4 of max(|g|, 1); the pdf to 1e-3 relative where it
    # is above 1e-6 (Gaussian tail values below that differ by their relative rounding)
    assert nerr.max() < 1e-4
    assert np.all(perr[rp > 1e-6] < 1e-3)
    assert np.all(np.abs(gp - rp)[rp <= 1e-6] < 1e-8)
    miss = np.nonzero(oc != po.HIT)[0]
    assert np.all(got["nee_pdf"][miss] == 0.0)
    return got, ref


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_flights_1d_renewal_plus_per_ray_length_scale_with_nee(precision):
    from gpis_extend_amd import workloads as W
    print(f"case: C3 65,536 1D flights + NEE, precision {precision}", flush=True)
    js = W.config3_json()
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=65536)
    _check_1d(js, rays, W.LIGHT_POS, f"C3 (65,536 1D flights + NEE, {precision})", precision=precision)


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_flights_1d_primary_rays_medium_length_scale(precision):
    from gpis_extend_amd import workloads as W
    js = W.config3_json()
    js["1D_sampling_scheme"] = "uni"
    med = _medium(js)
    rays = W.sphere_rays(med, n=16384)
    _check_1d(js, rays, None, f"1D primary (16,384 flights, {precision})", precision=precision)


def test_flights_1d_renewal_without_gradient():
    from gpis_extend_amd import workloads as W
    js = W.config3_json()
    js["correlation_context"] = "renewal"
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=16384, seed=99)
    _check_1d(js, rays, W.LIGHT_POS, "1D Renewal (16,384 flights + NEE)")


def test_flights_1d_renewal_plus_correlation_xy_with_nee():
    """1D_gradient_correlationXY: the transverse gradient at the endpoint and the NEE
    density follow the correlated law of seTransverseGradientMoments."""
    from gpis_extend_amd import workloads as W
    print("case: 1D Renewal+ correlationXY, 32,768 flights + NEE", flush=True)
    js = W.config3_json()
    js["1D_gradient_correlationXY"] = True
    med = _medium(js)
    rays = W.renewal1d_rays(med, n=32768, seed=7)
    got, ref = _check_1d(js, rays, W.LIGHT_POS, "1D Renewal+ correlationXY (32,768 flights + NEE)")
    oc = got["outcome"] & 0xFF
    same = oc == ref["outcome"]
    assert np.all(got["t"][same].view(np.uint32) == ref["t"][same].astype(np.float32).view(np.uint32))


def test_entry_points_reject_wrong_dimension():
    from gpis_extend_amd import GpisDevice, RayBatch, workloads as W
    from gpis_extend_amd._capi import GpisError
    med3 = _medium(W.config0_json())
    med1 = _medium(W.config3_json())
    r3 = RayBatch.from_numpy(W.sphere_rays(med3, n=64))
    d3 = GpisDevice(med3)
    with pytest.raises(GpisError, match="not 1D"):
        d3.trace_flights_1d(r3)
    d3.close()
    d1 = GpisDevice(med1)
    with pytest.raises(GpisError, match="gpis_trace_flights_1d"):
        d1.trace_flights(r3)
    d1.close()
