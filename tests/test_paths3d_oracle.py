"""The oracle's 3D path chaining (orc_trace_paths3d) pinned without the GPU: the bounce's
sine/cosine polynomial against libm, the diffuse direction's unit length and cosine law,
and each path's first and second flights against standalone oracle flights whose requests
are built here, in numpy, from the first flight's hit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_field, oracle_rays

BSEED = 7


def test_sincos_turn_matches_libm():
    u = np.concatenate([np.linspace(0, 1, 4097, endpoint=False, dtype=np.float32),
                        np.random.default_rng(3).random(4096, dtype=np.float32)])
    err = 0.0
    for x in u:
        s, c = po.sincos_turn(float(x))
        a = 2 * np.pi * float(x)
        err = max(err, abs(s - np.sin(a)), abs(c - np.cos(a)))
    assert err <= 4e-7, err
    # the four quarter turns are exact
    assert po.sincos_turn(0.0) == (0.0, 1.0)
    assert po.sincos_turn(0.25) == (1.0, 0.0) or po.sincos_turn(0.25) == (1.0, -0.0)
    s, c = po.sincos_turn(0.5)
    assert (s, c) in ((0.0, -1.0), (-0.0, -1.0))


@pytest.mark.parametrize("normal", [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0, -1, 0), (0.48, -0.6, 0.64),
                                    (-0.36, 0.48, -0.8)])
def test_diffuse_direction_cosine_law(normal):
    n = np.asarray(normal, np.float32)
    n = n / np.float32(np.linalg.norm(n))
    k = 6000
    cosv = np.zeros(k)
    jit = np.zeros(k)
    azi = np.zeros(k)
    # a tangent frame of our own to read the azimuth
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(n, a); t /= np.linalg.norm(t)
    b = np.cross(n, t)
    for i in range(k):
        d, j = po.diffuse_direction(i, 1 + i % 5, BSEED, n)
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1) < 1e-6
        cosv[i] = float(np.dot(d.astype(np.float64), n))
        jit[i] = j
        azi[i] = np.arctan2(np.dot(d, b), np.dot(d, t))
    assert cosv.min() >= -1e-6
    from scipy import stats
    # cosine-weighted hemisphere: cos^2 is uniform on [0, 1]; the azimuth is uniform
    assert stats.kstest(np.clip(cosv, 0, 1) ** 2, "uniform").pvalue > 1e-3
    assert stats.kstest((azi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    assert stats.kstest(jit, "uniform").pvalue > 1e-3
    assert abs(cosv.mean() - 2 / 3) < 0.02


def _box_exit(p, d, b=np.float32(1.0)):
    t = []
    for a in range(3):
        if d[a] != 0:
            t.append(max((b - p[a]) / d[a], (-b - p[a]) / d[a]))
        else:
            t.append(np.float32(3.4e38))
    return np.float32(max(min(t[0], min(t[1], t[2])), np.float32(0)))


def _unit_normal(g):
    g = g.astype(np.float32)
    l2 = g[0] * g[0] + g[1] * g[1]
    l2 = l2 + g[2] * g[2]
    inv = np.float32(1.0) / np.sqrt(l2)
    return (g * inv).astype(np.float32)


def _second_request(r0, fl, path, seeded):
    """The request of a path's second flight, built from its first flight's hit."""
    r = r0.copy()
    t = np.float32(fl["t"])
    o = r0["origin"].astype(np.float32)
    d = r0["dir"].astype(np.float32)
    p = (o + t * d).astype(np.float32)
    g = fl["grad"].astype(np.float32)
    w, jit = po.diffuse_direction(path, 1, BSEED, _unit_normal(g))
    r["origin"], r["dir"] = p, w
    r["near_t"] = 0.0
    r["far_t"] = _box_exit(p, w)
    r["jitter"] = jit
    r["first_scatter"] = 0
    r["last_val"] = 0.0
    r["last_aniso"] = g
    if seeded:
        r["pss"][3] = r0["pss"][3] + 1
    return r


@pytest.mark.parametrize("cfg", ["config1", "config2"])
def test_paths_chain_standalone_flights(cfg):
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W
    if cfg == "config1":
        med = SparseConvolutionNoiseMedium.from_json(W.config1_json())
        rays = W.camera_rays(med, width=64, height=64)
        rays = W.rays_subset(rays, np.arange(0, rays["ox"].shape[0], 8))
        seeded = False
    else:
        med = SparseConvolutionNoiseMedium.from_json(W.config2_json())
        rays = W.box_surface_rays(med, n=512)
        seeded = True
    F = oracle_field(med)
    orr = oracle_rays(rays, med)
    paths = F.trace_paths3d(orr, 1, BSEED, 1.0, seeded=seeded, nthreads=8)
    first = F.trace(orr, nthreads=8)
    assert np.array_equal(paths["outcome0"], first["outcome"])
    assert np.array_equal(paths["t0"], first["t"].astype(np.float32))
    assert np.array_equal(paths["grad0"], first["grad"])
    hit = np.nonzero(first["outcome"] == po.HIT)[0]
    assert hit.size > 20
    assert np.array_equal(paths["hits"] >= 1, first["outcome"] == po.HIT)
    seconds = np.zeros(hit.size, po.RAY_DTYPE)
    for k, i in enumerate(hit):
        seconds[k] = _second_request(orr[i], first[i], i, seeded)
    live = seconds["far_t"] > 0
    sec = F.trace(seconds[live], nthreads=8)
    ph = paths[hit]
    assert np.all(ph["flights"] == 2)
    assert np.array_equal(ph["t_last"][live], sec["t"].astype(np.float32))
    assert np.array_equal(ph["outcome_last"][live], sec["outcome"])
    assert np.array_equal(ph["hits"][live], 1 + (sec["outcome"] == po.HIT))
    # an empty second segment ends as an exit at farT
    assert np.all(ph["outcome_last"][~live] == po.EXIT)
    # paths without a first hit end after one flight
    miss = paths["hits"] == 0
    assert np.all(paths["flights"][miss] == 1)
    assert np.array_equal(paths["t_last"][miss], paths["t0"][miss])
