"""Single-precision 4x4 transforms as the reference's scene JSON builds them.

A procedural mean carries ``_invConfigTransform``, the float inverse of its "transform"
(GPFunctions.cpp:252-253).  The C ABI takes that inverse; these helpers compute it from
the JSON the way the reference does, in float32 with each operation rounded as written:

* ``from_json``  — JsonPtr::get(Mat4f &) (io/JsonPtr.cpp:108-183): a 16-element
  row-major array, or an object with "position", "scale" and "rotation" (degrees, the
  Y-X-Z Euler matrix of Mat4f::rotYXZ, Mat4f.cpp:118-130);
* ``invert``     — Mat4f::invert (Mat4f.hpp:75-101): the adjugate by Laplace expansion,
  det from the first row, scaled by the float reciprocal of det; identity when det is 0.

Sine and cosine come from the C library's sinf/cosf, as std::sin(float) does there.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util

import numpy as np

f32 = np.float32
_libm = C.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.restype = C.c_float
_libm.sinf.argtypes = [C.c_float]
_libm.cosf.restype = C.c_float
_libm.cosf.argtypes = [C.c_float]
_PI = f32(3.1415926536)


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32).reshape(16)


def rot_yxz(rot) -> np.ndarray:
    """Mat4f::rotYXZ: rotation by the Euler angles rot (degrees), row-major 4x4."""
    r = [f32(f32(f32(x) * _PI) / f32(180.0)) for x in rot]
    c = [f32(_libm.cosf(float(x))) for x in r]
    s = [f32(_libm.sinf(float(x))) for x in r]
    m = identity()
    m[0] = c[1] * c[2] - (s[1] * s[0]) * s[2]
    m[1] = (-c[1]) * s[2] - (s[1] * s[0]) * c[2]
    m[2] = (-s[1]) * c[0]
    m[4] = c[0] * s[2]
    m[5] = c[0] * c[2]
    m[6] = -s[0]
    m[8] = s[1] * c[2] + (c[1] * s[0]) * s[2]
    m[9] = (-s[1]) * s[2] + (c[1] * s[0]) * c[2]
    m[10] = c[1] * c[0]
    return m


def transform_point(m: np.ndarray, v) -> np.ndarray:
    """Mat4f * Vec3f: rows 0-2 of m applied to (v, 1), summed left to right."""
    v = [f32(x) for x in v]
    return np.array([((m[4 * i] * v[0] + m[4 * i + 1] * v[1]) + m[4 * i + 2] * v[2]) + m[4 * i + 3]
                     for i in range(3)], np.float32)


def _vec3(x) -> list:
    if isinstance(x, (list, tuple)):
        if len(x) != 3:
            raise ValueError(f"expected 3 components, got {len(x)}")
        return [f32(v) for v in x]
    return [f32(x)] * 3


def from_json(js) -> np.ndarray:
    """The row-major 4x4 float matrix of a JSON transform (array or object form)."""
    if js is None:
        return identity()
    if isinstance(js, (list, tuple)):
        if len(js) != 16:
            raise ValueError(f"a matrix needs 16 elements, received {len(js)}")
        return np.array([f32(v) for v in js], np.float32)
    for key in ("look_at", "up", "x_axis", "y_axis", "z_axis"):
        if key in js:
            raise ValueError(f"transform key {key!r} is not supported here; pass the 16-element matrix")
    axes = [np.array(a, np.float32) for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
    pos = _vec3(js.get("position", 0.0))
    if "scale" in js:
        sc = _vec3(js["scale"])
        axes = [axes[i] * sc[i] for i in range(3)]
    if "rotation" in js:
        t = rot_yxz(_vec3(js["rotation"]))
        axes = [transform_point(t, a) for a in axes]
    m = identity()
    for i in range(3):
        for j in range(3):
            m[4 * i + j] = axes[j][i]
        m[4 * i + 3] = pos[i]
    return m


def _cofactor(a: np.ndarray, r: int, c: int) -> np.float32:
    """Entry (r, c) of the adjugate: the signed minor of a without row c and column r,
    expanded along the minor's first column, products and sums left to right."""
    i1, i2, i3 = [i for i in range(4) if i != c]
    j1, j2, j3 = [j for j in range(4) if j != r]
    e = lambda i, j: a[4 * i + j]
    terms = [(1, e(i1, j1), e(i2, j2), e(i3, j3)), (-1, e(i1, j1), e(i2, j3), e(i3, j2)),
             (-1, e(i2, j1), e(i1, j2), e(i3, j3)), (1, e(i2, j1), e(i1, j3), e(i3, j2)),
             (1, e(i3, j1), e(i1, j2), e(i2, j3)), (-1, e(i3, j1), e(i1, j3), e(i2, j2))]
    sign = 1 if (r + c) % 2 == 0 else -1
    acc = None
    for s, x, y, z in terms:
        t = f32(f32(x * y) * z)
        if s * sign < 0:
            t = -t
        acc = t if acc is None else f32(acc + t)
    return acc


def invert(m) -> np.ndarray:
    """Mat4f::invert of a row-major 4x4 float matrix."""
    a = np.asarray(m, np.float32).reshape(16)
    inv = np.array([_cofactor(a, k // 4, k % 4) for k in range(16)], np.float32)
    det = f32(f32(f32(a[0] * inv[0] + a[1] * inv[4]) + a[2] * inv[8]) + a[3] * inv[12])
    if det == 0.0:
        return identity()
    inv_det = f32(f32(1.0) / det)
    return (inv * inv_det).astype(np.float32)
