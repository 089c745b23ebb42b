"""Writes tests/golden/sdf_means.npz from the reference's own SDF code
(SdfFunctions::evaluateJet and the float API of SdfFunctionsQueries.cpp, Mat4f.cpp,
compiled into oracle/_ref/libgpis_ref_leaf.so by oracle/build_ref.sh): jets of every
function with queries, the float API of the others, ProceduralMean::evaluateJet's
products on the reference's Eigen, and Mat4f::rotYXZ / invert.  Run in the build
container, where the reference tree exists; the committed .npz travels instead.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sdf_means.npz")
QUERY_FNS = (po.SDF_KNOB, po.SDF_TWO_SPHERES, po.SDF_ROUNDED_U_GROOVE, po.SDF_TWO_SLABS, po.SDF_THREE_SLABS)
FLOAT_FNS = (po.SDF_KNOB_INNER, po.SDF_KNOB_OUTER, po.SDF_PLANE)
SCALES = {po.SDF_KNOB: 1.5, po.SDF_TWO_SPHERES: 25.0, po.SDF_ROUNDED_U_GROOVE: 2.0, po.SDF_TWO_SLABS: 1.0,
          po.SDF_THREE_SLABS: 1.0, po.SDF_KNOB_INNER: 1.5, po.SDF_KNOB_OUTER: 1.5, po.SDF_PLANE: 2.0}
# (rotation degrees, uniform scale, position, min, scale, offset) of the procedural means
MEANS = [((0, 0, 0), 1.0, (0.0, 0.8, 0.0), -3.4028234663852886e38, 1.0, 0.0),
         ((30, 45, 10), 0.7, (0.1, 0.2, -0.3), -0.05, 0.5, 0.01),
         ((0, 0, 0), 0.1, (0.0, 0.0, 0.0), -3.4028234663852886e38, 0.1, 0.0),
         ((0, 90, 0), 2.0, (0.0, 0.0, 0.0), 0.0, 0.0, 0.25)]


def seam_points(rng, n, scale):
    """Uniform points with exact zeros, mirrored and rounded coordinates mixed in, so
    that the CSG's ties, folds and zero branches are exercised."""
    p = rng.uniform(-scale, scale, (n, 3))
    p[::7, 1] = 0.0
    p[::11, 0] = 0.0
    p[::13, 2] = -p[::13, 2]
    p[::17] = np.round(p[::17], 1)
    return p


def ref_lib():
    L = C.CDLL(po.REF_PATH)
    d3, f3, f16 = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_float)
    L.ref_sdf_jet.restype = C.c_double
    L.ref_sdf_jet.argtypes = [C.c_int, d3, d3]
    L.ref_sdf_eval.restype = C.c_float
    L.ref_sdf_eval.argtypes = [C.c_int, f3]
    L.ref_proc_mean_jet.restype = C.c_double
    L.ref_proc_mean_jet.argtypes = [C.c_int, f16, C.c_float, C.c_float, C.c_float, d3, d3]
    L.ref_mat4f_rot_yxz.argtypes = [f3, f16]
    L.ref_mat4f_invert.argtypes = [f16, f16]
    L.ref_mat4f_transform_vector.argtypes = [f16, f3, f3]
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def ref_inverse(L, rot, scale, pos):
    """The reference's _invConfigTransform of {"position", "scale", "rotation"}: columns
    rotYXZ(rot) * (scale * axis), translation pos, inverted by Mat4f::invert."""
    R = np.zeros(16, np.float32)
    L.ref_mat4f_rot_yxz(_p(np.array(rot, np.float32), C.c_float), _p(R, C.c_float))
    M = np.eye(4, dtype=np.float32).reshape(16)
    for j, ax in enumerate(np.eye(3, dtype=np.float32)):
        out = np.zeros(3, np.float32)
        L.ref_mat4f_transform_vector(_p(R, C.c_float), _p(ax * np.float32(scale), C.c_float), _p(out, C.c_float))
        for i in range(3):
            M[4 * i + j] = out[i]
    M[[3, 7, 11]] = np.array(pos, np.float32)
    inv = np.zeros(16, np.float32)
    L.ref_mat4f_invert(_p(M, C.c_float), _p(inv, C.c_float))
    return M, inv


def main():
    assert po.ref_available(), "build oracle/_ref first (oracle/build_ref.sh)"
    L = ref_lib()
    rng = np.random.default_rng(20261203)
    out = {}
    for fn in QUERY_FNS:
        p = seam_points(rng, 512, SCALES[fn])
        vg = np.zeros((len(p), 4))
        for k in range(len(p)):
            g = np.zeros(3)
            vg[k, 0] = L.ref_sdf_jet(fn, _p(np.ascontiguousarray(p[k]), C.c_double), _p(g, C.c_double))
            vg[k, 1:] = g
        out[f"jet{fn}_p"], out[f"jet{fn}_vg"] = p, vg
    for fn in FLOAT_FNS:
        p = seam_points(rng, 512, SCALES[fn]).astype(np.float32)
        out[f"eval{fn}_p"] = p
        out[f"eval{fn}_v"] = np.array([L.ref_sdf_eval(fn, _p(np.ascontiguousarray(q), C.c_float)) for q in p],
                                      np.float32)
    for mi, (rot, sc, pos, mn, s, off) in enumerate(MEANS):
        M, inv = ref_inverse(L, rot, sc, pos)
        out[f"mean{mi}_forward"], out[f"mean{mi}_inverse"] = M, inv
        for fn in QUERY_FNS:
            p = seam_points(rng, 128, 2.0)
            vg = np.zeros((len(p), 4))
            for k in range(len(p)):
                g = np.zeros(3)
                vg[k, 0] = L.ref_proc_mean_jet(fn, _p(inv, C.c_float), mn, s, off,
                                               _p(np.ascontiguousarray(p[k]), C.c_double), _p(g, C.c_double))
                vg[k, 1:] = g
            out[f"mean{mi}_fn{fn}_p"], out[f"mean{mi}_fn{fn}_vg"] = p, vg
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, len(out), "arrays")


if __name__ == "__main__":
    main()
