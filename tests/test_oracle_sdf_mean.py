"""The procedural SDF means (ProceduralMean over SdfFunctions, GPFunctions.cpp:250-298,
ProceduralMeanSdf.cpp:32-38, SdfFunctionsQueries.cpp): the oracle's restatement pinned
to the bit against golden vectors written from the reference's own SDF code
(tests/golden/sdf_means.npz, by tests/golden/make_sdf_golden.py) and, where the
reference build is present, against it live; closed forms of the simple shapes;
finite differences of the knob; the mean's floor, scale and forward-difference rules;
the host's transform helpers and JSON parsing."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "sdf_means.npz"))
QUERY_FNS = (po.SDF_KNOB, po.SDF_TWO_SPHERES, po.SDF_ROUNDED_U_GROOVE, po.SDF_TWO_SLABS, po.SDF_THREE_SLABS)
FLOAT_FNS = (po.SDF_KNOB_INNER, po.SDF_KNOB_OUTER, po.SDF_PLANE)


def _o():
    L = po.lib()
    d3 = C.POINTER(C.c_double)
    L.orc_sdf_jet.restype = C.c_double
    L.orc_sdf_jet.argtypes = [C.c_int, d3, d3]
    L.orc_sdf_eval.restype = C.c_float
    L.orc_sdf_eval.argtypes = [C.c_int, C.POINTER(C.c_float)]
    L.orc_sdf_mean.restype = C.c_double
    L.orc_sdf_mean.argtypes = [C.POINTER(po.MediumDesc), d3, d3]
    return L


def jet(fn, p):
    L = _o()
    out = np.zeros((len(p), 4))
    for k, q in enumerate(np.asarray(p, np.float64)):
        g = np.zeros(3)
        q = np.ascontiguousarray(q)
        out[k, 0] = L.orc_sdf_jet(fn, q.ctypes.data_as(C.POINTER(C.c_double)), g.ctypes.data_as(C.POINTER(C.c_double)))
        out[k, 1:] = g
    return out


def eval_float(fn, p):
    L = _o()
    return np.array([L.orc_sdf_eval(fn, np.ascontiguousarray(q, np.float32).ctypes.data_as(C.POINTER(C.c_float)))
                     for q in p], np.float32)


def mean_desc(fn, inverse12, minimum=-3.4028234663852886e38, scale=1.0, offset=0.0):
    return po.make_desc(mean_kind=po.MEAN_PROC, sdf_function=fn, sdf_inverse_transform=list(inverse12),
                        sdf_min=minimum, sdf_scale=scale, sdf_offset=offset)


def mean(desc, p):
    L = _o()
    out = np.zeros((len(p), 4))
    for k, q in enumerate(np.asarray(p, np.float64)):
        g = np.zeros(3)
        q = np.ascontiguousarray(q)
        out[k, 0] = L.orc_sdf_mean(C.byref(desc), q.ctypes.data_as(C.POINTER(C.c_double)),
                                   g.ctypes.data_as(C.POINTER(C.c_double)))
        out[k, 1:] = g
    return out


@pytest.mark.parametrize("fn", QUERY_FNS)
def test_jets_equal_the_reference_golden_vectors(fn):
    got = jet(fn, GOLD[f"jet{fn}_p"])
    assert got.tobytes() == GOLD[f"jet{fn}_vg"].tobytes()


@pytest.mark.parametrize("fn", FLOAT_FNS)
def test_float_api_equals_the_reference_golden_vectors(fn):
    got = eval_float(fn, GOLD[f"eval{fn}_p"])
    assert got.tobytes() == GOLD[f"eval{fn}_v"].tobytes()


@pytest.mark.parametrize("mi", range(4))
def test_procedural_mean_equals_the_reference_golden_vectors(mi):
    from tests.golden.make_sdf_golden import MEANS
    rot, sc, pos, mn, s, off = MEANS[mi]
    inv = GOLD[f"mean{mi}_inverse"]
    for fn in QUERY_FNS:
        d = mean_desc(fn, inv[:12], mn, s, off)
        got = mean(d, GOLD[f"mean{mi}_fn{fn}_p"])
        assert got.tobytes() == GOLD[f"mean{mi}_fn{fn}_vg"].tobytes(), (mi, fn)


@pytest.mark.parametrize("mi", range(4))
def test_host_transform_helpers_equal_the_reference(mi):
    from gpis_extend_amd import mat4f
    from tests.golden.make_sdf_golden import MEANS
    rot, sc, pos = MEANS[mi][:3]
    fwd = mat4f.from_json({"position": list(pos), "scale": sc, "rotation": list(rot)})
    assert fwd.tobytes() == GOLD[f"mean{mi}_forward"].tobytes()
    assert mat4f.invert(fwd).tobytes() == GOLD[f"mean{mi}_inverse"].tobytes()


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent (GPU box)")
def test_live_against_the_reference_build():
    from tests.golden import make_sdf_golden as G
    L = G.ref_lib()
    rng = np.random.default_rng(7)
    for fn in QUERY_FNS:
        p = G.seam_points(rng, 2000, G.SCALES[fn])
        ref = np.zeros((len(p), 4))
        for k in range(len(p)):
            g = np.zeros(3)
            ref[k, 0] = L.ref_sdf_jet(fn, G._p(np.ascontiguousarray(p[k]), C.c_double), G._p(g, C.c_double))
            ref[k, 1:] = g
        assert jet(fn, p).tobytes() == ref.tobytes(), fn
    from gpis_extend_amd import mat4f
    for k in range(200):
        rot, sc, pos = rng.uniform(-180, 180, 3), rng.uniform(0.05, 3.0), rng.uniform(-1, 1, 3)
        fwd, inv = G.ref_inverse(L, rot.astype(np.float32), np.float32(sc), pos.astype(np.float32))
        mine = mat4f.from_json({"position": [float(np.float32(v)) for v in pos], "scale": float(np.float32(sc)),
                                "rotation": [float(np.float32(v)) for v in rot]})
        assert mine.tobytes() == fwd.tobytes()
        assert mat4f.invert(mine).tobytes() == inv.tobytes()


def test_simple_shapes_match_their_closed_forms():
    rng = np.random.default_rng(11)
    p = rng.uniform(-25, 25, (3000, 3))
    # two spheres of radius 9.5 at y = +-10: distance to the nearer centre minus 9.5
    v = jet(po.SDF_TWO_SPHERES, p)
    c = np.stack([p[:, 0], np.abs(p[:, 1]) - 10.0, p[:, 2]], 1)
    r = np.linalg.norm(c, axis=1)
    assert np.allclose(v[:, 0], r - 9.5, rtol=0, atol=1e-12)
    g = c / r[:, None] * np.where(np.arange(3) == 1, np.sign(p[:, 1])[:, None], 1.0)
    assert np.allclose(v[:, 1:], g, atol=1e-12)
    # slabs: exact max/min of face distances (pure subtraction, max, min: equal to the bit)
    p = rng.uniform(-1, 1, (3000, 3))
    slab = lambda q: np.maximum(np.maximum(np.abs(q[:, 0]) - 0.6, np.abs(q[:, 1]) - 0.01), np.abs(q[:, 2]) - 0.6)
    two = np.minimum(slab(p + [-0.18, -0.02, 0.0]), slab(p + [0.18, 0.02, 0.0]))
    assert np.array_equal(jet(po.SDF_TWO_SLABS, p)[:, 0], two)
    assert np.array_equal(jet(po.SDF_THREE_SLABS, p)[:, 0], np.minimum(two, slab(p + [-0.18, 0.06, 0.0])))
    # the rounded U channel: 1 - |(x, min(y, 0))|
    p = rng.uniform(-2, 2, (3000, 3))
    u = jet(po.SDF_ROUNDED_U_GROOVE, p)
    assert np.allclose(u[:, 0], 1.0 - np.hypot(p[:, 0], np.minimum(p[:, 1], 0.0)), atol=1e-15)
    # knob_inner is a sphere of radius 0.6 (0.75 scaled by 0.8); plane is y
    p32 = rng.uniform(-1, 1, (2000, 3)).astype(np.float32)
    assert np.allclose(eval_float(po.SDF_KNOB_INNER, p32), np.linalg.norm(p32.astype(np.float64), axis=1) - 0.6,
                       atol=2e-7)
    assert np.array_equal(eval_float(po.SDF_PLANE, p32), p32[:, 1])


def test_knob_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.3, 1.3, (400, 3))
    v = jet(po.SDF_KNOB, p)
    h = 1e-6
    fd = np.stack([(jet(po.SDF_KNOB, p + h * e)[:, 0] - jet(po.SDF_KNOB, p - h * e)[:, 0]) / (2 * h)
                   for e in np.eye(3)], 1)
    err = np.linalg.norm(fd - v[:, 1:], axis=1)
    # CSG seams make a few points non-differentiable; the rest agree to FD accuracy
    assert np.mean(err < 1e-5) > 0.97
    # a signed distance: unit gradient almost everywhere (smooth unions and the 0.7 base scale
    # aside), up to the float size factors double(0.8f) * float(1 / 0.8f) = 1 + 1.5e-8
    # (the base is a distance scaled by 0.7)
    n = np.linalg.norm(v[:, 1:], axis=1)
    unit = (np.abs(n - 1.0) < 1e-7) | (np.abs(n - 0.7) < 1e-7)
    assert np.mean(unit) > 0.85  # the cones (slope pi/3) and smooth unions are the rest


def test_mean_floor_scale_and_forward_differences():
    ident = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)
    p = np.random.default_rng(2).uniform(-1, 1, (200, 3))
    base = jet(po.SDF_TWO_SLABS, p)
    # scale 0: the value floors at max(offset, min) and the gradient is zero
    m = mean(mean_desc(po.SDF_TWO_SLABS, ident, -1.0, 0.0, 0.25), p)
    assert np.all(m[:, 0] == 0.25) and np.all(m[:, 1:] == 0.0)
    # min: values below it are lifted to it, with a zero gradient
    m = mean(mean_desc(po.SDF_TWO_SLABS, ident, 0.0, 2.0, 0.0), p)
    below = 2.0 * base[:, 0] < 0.0
    assert np.all(m[below, 0] == 0.0) and np.all(m[below, 1:] == 0.0)
    assert np.array_equal(m[~below, 0], 2.0 * base[~below, 0])
    assert np.array_equal(m[~below, 1:], 2.0 * base[~below, 1:])
    # a function without queries: float API through the float inverse transform, gradient
    # by forward differences with eps = 0.001f; the plane scaled by 1/4 has slope 4 * scale
    inv = (4, 0, 0, 0, 0, 4, 0, 0, 0, 0, 4, 0)
    m = mean(mean_desc(po.SDF_PLANE, inv, -3.4028234663852886e38, 0.5, 0.0), p)
    assert np.array_equal(m[:, 0], (np.float32(4) * p[:, 1].astype(np.float32) * np.float32(0.5)).astype(np.float64))
    assert np.allclose(m[:, 2], 2.0, atol=1e-3) and np.allclose(m[:, [1, 3]], 0.0, atol=1e-3)


def test_medium_json_procedural_means():
    from gpis_extend_amd import SparseConvolutionNoiseMedium, workloads as W, _capi as A
    from gpis_extend_amd.medium import MediumConfigError
    # the two reference scenes' means (two-spheres-backlight/scene.json, classic-global/shaderball-*.json)
    js = W.with_procedural_mean(W.config0_json(), "two_spheres", {"scale": 0.1}, scale=0.1)
    m = SparseConvolutionNoiseMedium.from_json(js)
    assert m.mean_kind == A.MEAN_PROCEDURAL and m.sdf_function == po.SDF_TWO_SPHERES
    assert m.sdf_inverse_transform == (10.0, 0.0, 0.0, 0.0, 0.0, 10.0, 0.0, 0.0, 0.0, 0.0, 10.0, 0.0)
    assert np.float32(m.sdf_scale) == np.float32(0.1) and m.sdf_min == -3.4028234663852886e38
    js = W.with_procedural_mean(W.config0_json(), "knob", {"position": [0.0, 0.8, 0.0], "scale": 1.0,
                                                           "rotation": [0.0, 0.0, 0.0]})
    m = SparseConvolutionNoiseMedium.from_json(js)
    assert m.sdf_function == po.SDF_KNOB
    assert np.allclose(m.sdf_inverse_transform, (1, 0, 0, 0, 0, 1, 0, -0.8, 0, 0, 1, 0), atol=1e-7)
    d = m.to_desc()
    assert d.mean_kind == A.MEAN_PROCEDURAL and d.sdf_function == po.SDF_KNOB
    for bad in ({"type": "procedural", "func": "teapot"},
                {"type": "procedural", "f": {"type": "constant", "value": 1.0}},
                {"type": "procedural", "func": "knob", "transform": {"look_at": [0, 0, 1]}}):
        js = W.config0_json()
        js["gaussian_process"]["mean"] = bad
        with pytest.raises(MediumConfigError):
            SparseConvolutionNoiseMedium.from_json(js)
