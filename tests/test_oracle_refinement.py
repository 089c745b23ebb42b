"""The oracle's crossing refinement (finishGpisSegment's secant loop,
SparseConvolutionNoiseMedium.cpp:308-330) pinned by independent evaluations of the field:
for every crossing flight, the lattice is re-scanned here, the bracket [a, b] = [t_k - step,
t_k] and its values (pf, fc) found, and the candidates t_j = lerp(a, b, fac_j) with
fac_0 = pf / (pf - fc), fac_{j+1} = 0.9 fac_j formed; their signs, from the field evaluated
here, must make the oracle's distance the rule's answer:
  * a secant point on the origin's side is returned as it is (t = t_0);
  * otherwise the returned t_j lies beyond the crossing (sign opposite to the origin's) and
    so do t_0 .. t_j, while t_{j+1} is back on the origin's side (the previous candidate
    is returned, not the one that crossed back);
  * t = 0 exactly when every candidate with fac > 0.01 lies beyond the crossing.
A different shrink factor or returning the last candidate instead of the previous one
fails these checks."""
import numpy as np

from oracle import pyoracle as po
from tests.helpers import oracle_desc, oracle_field, oracle_rays


def _points(rec, ts):
    o = np.asarray(rec["origin"], np.float32).astype(np.float64)
    d = np.asarray(rec["dir"], np.float32).astype(np.float64)
    return (o[None, :] + np.asarray(ts, np.float64)[:, None] * d[None, :]).astype(np.float32)


def _sign(v):
    return -1 if v < 0 else 1


def _check(rec, res, fvals, step_size, min_step, stats):
    """fvals(ts) -> float values of the flight's realization at the points of ts."""
    near, far, jit = np.float32(rec["near_t"]), np.float32(rec["far_t"]), np.float32(rec["jitter"])
    step = np.float32((far - near) / np.float32(min_step))
    if np.float32(step_size) < step:
        step = np.float32(step_size)
    t0 = float(np.float32(near + np.float32(step * jit)))
    n = int(np.ceil((float(far) - t0) / float(step))) + 2
    lat = t0 + np.arange(n) * float(step)
    lat = lat[lat < float(far)]
    f_near = float(fvals([float(near)])[0])
    fl = fvals(lat).astype(np.float64)
    first = bool(rec["first_scatter"])
    sign0 = _sign(f_near) if first else _sign(fl[0])
    k = next((i for i in range(0 if first else 1, len(lat)) if _sign(fl[i]) != sign0), None)
    assert k is not None, "the oracle reports a crossing the lattice does not have"
    pf = f_near if k == 0 else fl[k - 1]
    fc = fl[k]
    a, b = lat[k] - float(step), lat[k]
    fac = [pf / (pf - fc)]
    while fac[-1] * 0.9 > 0.01:
        fac.append(fac[-1] * 0.9)
    cand = np.array([a * (1.0 - x) + b * x for x in fac])
    sg = [_sign(v) for v in fvals(cand)]
    t = float(res["t"])
    if _sign(sg[0]) == sign0:
        assert t == cand[0], (t, cand[0])
        stats["short"] += 1
        return
    beyond = [s != sign0 for s in sg]
    if t == 0.0:
        assert all(beyond), "t = 0 although a candidate above the 0.01 cut-off is on the origin's side"
        stats["cutoff"] += 1
        return
    hits = np.nonzero(cand == t)[0]
    assert hits.size == 1, (t, cand[:4])
    j = int(hits[0])
    assert all(beyond[:j + 1]), "a candidate before the returned one is back on the origin's side"
    assert j + 1 < len(cand) and not beyond[j + 1], "the candidate after the returned one is not back"
    stats["shrunk" if j > 0 else "first_beyond"] += 1


def _stats():
    return {"short": 0, "first_beyond": 0, "shrunk": 0, "cutoff": 0}


def test_refinement_on_stationary_flights():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    med = SparseConvolutionNoiseMedium.from_json(W.config0_json())
    F = oracle_field(med)
    rays = W.sphere_rays(med, n=3000)
    orr = oracle_rays(rays, med)
    res = F.trace(orr, nthreads=8)
    stats = _stats()
    for i in np.nonzero(res["outcome"] != po.EXIT)[0]:
        _check(orr[i], res[i], lambda ts: F.eval_points(_points(orr[i], ts), med.seed)[:, 0],
               med.step_size, med.min_step, stats)
    print(stats)
    assert stats["short"] >= 20 and stats["shrunk"] >= 20 and stats["first_beyond"] >= 20


def test_refinement_on_conditioned_multires_flights():
    """Surface-born Renewal+ flights of the nonstationary workload (conditioned on (0, the
    face normal), two-level field): the candidates evaluated on each flight's own
    conditioned realization; these flights also reach the 0.01 cut-off."""
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    med = SparseConvolutionNoiseMedium.from_json(W.config2_json())
    F = oracle_field(med)
    rays = W.box_surface_rays(med, n=1500)
    orr = oracle_rays(rays, med)
    res = F.trace(orr, nthreads=8)
    stats = _stats()
    done = 0
    for i in np.nonzero((res["outcome"] != po.EXIT) & (res["steps"] > 0))[0]:
        rec = orr[i]
        _check(rec, res[i], lambda ts: np.array([F.conditioned_eval(rec, q)[0] for q in _points(rec, ts)]),
               med.step_size, med.min_step, stats)
        done += 1
    print(done, stats)
    assert stats["short"] >= 10 and stats["shrunk"] >= 10 and stats["cutoff"] >= 1
