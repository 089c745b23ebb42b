// gpis_sdf.hpp — the procedural signed-distance means of the sparse_conv_noise medium:
// ProceduralMean over an SdfFunctions function (GPFunctions.cpp:250-298,
// ProceduralMeanSdf.cpp:11-38, SdfFunctions.hpp:13-91, SdfFunctionsQueries.cpp:14-466).
//
// Every shape is one constructive-solid-geometry program written once as a template over
// its number type, and run with two of them, as the reference does:
//   * double — the float API (SdfFunctions::knob_inner, knob_outer: knobGeometry on
//     doubles, rounded to float), used by the mean of the functions without queries;
//   * Jet — a value with its gradient and the two structural flags of the reference's
//     point queries (constant, affine), used by the functions with queries (knob,
//     two_spheres, two_slabs, three_slabs; rounded_u_groove has a closed form).
// The flags only decide branches at exact zeros and in the round-off band of a CSG
// minimum; they are carried so that those cases resolve as in the reference.
//
// Arithmetic is double, unfused (#pragma clang fp contract(off)), in the reference's order.
// The three-argument hypot is the C++17 std::hypot of libstdc++ 11 (scale by the largest
// magnitude, then sqrt of the sum of the three squared ratios, left to right).
// The SDF shapes below (the knob, two-spheres, slab and groove programs) follow the shape
// formulas of the reference's src/core/math/SdfFunctions.cpp, which carries these notices:
//
//   Copyright 2020 Towaki Takikawa @yongyuanxi
//   The MIT License
//
//   The MIT License (MIT)
//   Copyright (c) 2021, NVIDIA CORPORATION.
//   Permission is hereby granted, free of charge, to any person obtaining a copy of
//   this software and associated documentation files (the "Software"), to deal in
//   the Software without restriction, including without limitation the rights to
//   use, copy, modify, merge, publish, distribute, sublicense, and/or sell copies of
//   the Software, and to permit persons to whom the Software is furnished to do so,
//   subject to the following conditions:
//   The above copyright notice and this permission notice shall be included in all
//   copies or substantial portions of the Software.
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
//   IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY, FITNESS
//   FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE AUTHORS OR
//   COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER
//   IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN
//   CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
//
#pragma once

#ifndef GPIS_HOST_CHECK
#include <hip/hip_runtime.h>
#endif

#include <cmath>

#define GPIS_HD __host__ __device__ __forceinline__

namespace gpis {
namespace sdf {

// SdfFunctions::Function (SdfFunctions.hpp:13-22), in its order.
enum Fn { KNOB = 0, KNOB_INNER = 1, KNOB_OUTER = 2, TWO_SPHERES = 3, PLANE = 4, ROUNDED_U_GROOVE = 5,
          TWO_SLABS = 6, THREE_SLABS = 7, FN_COUNT = 8 };

// SdfFunctions::supportsQueries (SdfFunctionsQueries.cpp:395-398).
GPIS_HD bool supports_queries(int fn) {
    return fn == KNOB || fn == ROUNDED_U_GROOVE || fn == TWO_SPHERES || fn == TWO_SLABS || fn == THREE_SLABS;
}

// Parameters of one procedural mean, resident in device memory.
struct MeanParams {
    double w[9];        // world -> local 3x3 (rows of the inverse configuration transform), row-major
    double t[3];        // world -> local translation
    float inv[12];      // the inverse configuration transform's rows 0-2 as floats (the float API)
    float rot_p90[9];   // Mat4f::rotAxis((1,0,0), +90), row-major 3x3
    float rot_m90[9];   // ... -90
    float rot_m45[9];   // ... -45
    float minv, scale, offset;
    int fn;
    int queries;        // supports_queries(fn)
};

// ---------------------------------------------------------------- scalars --
GPIS_HD double hypot3(double x, double y, double z) {
#pragma clang fp contract(off)
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != 0.0 || a != a) {
        const double u = x / a, v = y / a, w = z / a;
        return a * sqrt(u * u + v * v + w * w);
    }
    return 0.0;
}
// The round-off guard of the point queries: 128 ulp(1) * (1 + |v|).
GPIS_HD double guard(double v) {
#pragma clang fp contract(off)
    return 0x1p-45 * (1.0 + fabs(v));
}

// ------------------------------------------------------------ plain double --
GPIS_HD double cst(double x, double) { return x; }
GPIS_HD double add(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
GPIS_HD double mul(double a, double s) {
#pragma clang fp contract(off)
    return a * s;
}
GPIS_HD double neg(double a) { return -a; }
GPIS_HD double sub(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
GPIS_HD double minimum(double a, double b) { return b < a ? b : a; }  // std::min
GPIS_HD double maximum(double a, double b) { return a < b ? b : a; }  // std::max
GPIS_HD double norm(double a, double b, double c) { return hypot3(a, b, c); }
GPIS_HD double smooth_union(double a, double b, double k) {
#pragma clang fp contract(off)
    double h = 0.5 + 0.5 * (b - a) / k;
    h = h < 0.0 ? 0.0 : (1.0 < h ? 1.0 : h);
    return b + (a - b) * h - k * h * (1.0 - h);
}

// -------------------------------------------------------------- point jets --
struct Jet {
    double v, g0, g1, g2;
    bool constant, affine;
};
GPIS_HD Jet cst(double x, Jet) { return Jet{x, 0.0, 0.0, 0.0, true, true}; }
GPIS_HD Jet coord(double x, int axis) {
    return Jet{x, axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0, axis == 2 ? 1.0 : 0.0, false, true};
}
GPIS_HD Jet add(const Jet &a, const Jet &b) {
#pragma clang fp contract(off)
    return Jet{a.v + b.v, a.g0 + b.g0, a.g1 + b.g1, a.g2 + b.g2, a.constant && b.constant, a.affine && b.affine};
}
GPIS_HD Jet mul(const Jet &a, double s) {
#pragma clang fp contract(off)
    if (s == 0.0) return Jet{0.0, 0.0, 0.0, 0.0, true, true};
    return Jet{a.v * s, a.g0 * s, a.g1 * s, a.g2 * s, a.constant, a.affine};
}
GPIS_HD Jet neg(const Jet &a) { return mul(a, -1.0); }
GPIS_HD Jet sub(const Jet &a, const Jet &b) { return add(a, neg(b)); }
GPIS_HD Jet minimum(const Jet &a, const Jet &b) {
#pragma clang fp contract(off)
    if (a.constant && b.constant) return Jet{b.v < a.v ? b.v : a.v, 0.0, 0.0, 0.0, true, true};
    if (a.v + guard(a.v) < b.v - guard(b.v)) return a;
    if (b.v + guard(b.v) < a.v - guard(a.v)) return b;
    Jet c = a.v <= b.v ? a : b;
    c.constant = false;
    c.affine = false;
    return c;
}
GPIS_HD Jet maximum(const Jet &a, const Jet &b) { return neg(minimum(neg(a), neg(b))); }
GPIS_HD bool is_zero(const Jet &a) { return a.v == 0.0 && a.constant; }
GPIS_HD Jet norm(const Jet &a, const Jet &b, const Jet &c) {
#pragma clang fp contract(off)
    if (is_zero(b) && is_zero(c)) return maximum(a, neg(a));
    if (is_zero(a) && is_zero(c)) return maximum(b, neg(b));
    if (is_zero(a) && is_zero(b)) return maximum(c, neg(c));
    Jet r{hypot3(a.v, b.v, c.v), 0.0, 0.0, 0.0, true, true};
    if (r.v > 0.0) {
        const double ra = a.v / r.v, rb = b.v / r.v, rc = c.v / r.v;
        r.g0 = r.g0 + a.g0 * ra; r.g1 = r.g1 + a.g1 * ra; r.g2 = r.g2 + a.g2 * ra;
        r.g0 = r.g0 + b.g0 * rb; r.g1 = r.g1 + b.g1 * rb; r.g2 = r.g2 + b.g2 * rb;
        r.g0 = r.g0 + c.g0 * rc; r.g1 = r.g1 + c.g1 * rc; r.g2 = r.g2 + c.g2 * rc;
    }
    r.constant = a.constant && b.constant && c.constant;
    if (a.affine && b.affine && c.affine)
        r.constant = a.g0 == 0.0 && a.g1 == 0.0 && a.g2 == 0.0 && b.g0 == 0.0 && b.g1 == 0.0 && b.g2 == 0.0 &&
                     c.g0 == 0.0 && c.g1 == 0.0 && c.g2 == 0.0;
    r.affine = r.constant;
    return r;
}
GPIS_HD Jet smooth_union(const Jet &a, const Jet &b, double k) {
#pragma clang fp contract(off)
    const double d = b.v - a.v, err = guard(a.v) + guard(b.v);
    if (d - err >= k) return a;
    if (d + err <= -k) return b;
    double h = 0.5 + 0.5 * d / k;
    h = h < 0.0 ? 0.0 : (1.0 < h ? 1.0 : h);
    const double h1 = 1.0 - h;
    Jet c{b.v + (a.v - b.v) * h - k * h * h1, h * a.g0 + h1 * b.g0, h * a.g1 + h1 * b.g1, h * a.g2 + h1 * b.g2,
          a.constant && b.constant, false};
    c.affine = c.constant;
    return c;
}

// ------------------------------------------------------ shared CSG program --
template <class T>
struct P3 {
    T x, y, z;
};
template <class T>
GPIS_HD T subtract(const T &a, const T &b, double k) { return neg(smooth_union(a, neg(b), k)); }
template <class T>
GPIS_HD T absval(const T &a) { return maximum(a, neg(a)); }
template <class T>
GPIS_HD P3<T> shifted(const P3<T> &p, double x, double y, double z) {
    const T o = p.x;
    return P3<T>{add(p.x, cst(x, o)), add(p.y, cst(y, o)), add(p.z, cst(z, o))};
}
template <class T>
GPIS_HD P3<T> rotated(const P3<T> &p, const float *m) {
    return P3<T>{add(add(mul(p.x, (double)m[0]), mul(p.y, (double)m[1])), mul(p.z, (double)m[2])),
                 add(add(mul(p.x, (double)m[3]), mul(p.y, (double)m[4])), mul(p.z, (double)m[5])),
                 add(add(mul(p.x, (double)m[6]), mul(p.y, (double)m[7])), mul(p.z, (double)m[8]))};
}
template <class T>
GPIS_HD T sphere(const P3<T> &p, double r) { return sub(norm(p.x, p.y, p.z), cst(r, p.x)); }
template <class T>
GPIS_HD T cone(const P3<T> &p) {
    return mul(add(norm(p.x, p.y, cst(0.0, p.x)), p.z), (double)(3.1415926536f / 3.f));
}
template <class T>
GPIS_HD T cylinder(const P3<T> &p, double h, double r) {
    const T zero = cst(0.0, p.x);
    const T x = sub(norm(p.x, p.z, zero), cst(h, p.x));
    const T y = sub(absval(p.y), cst(r, p.x));
    return add(minimum(maximum(x, y), zero), norm(maximum(x, zero), maximum(y, zero), zero));
}
template <class T>
GPIS_HD T prism(const P3<T> &p) {
    return maximum(sub(absval(p.z), cst((double)0.3f, p.x)),
                   sub(maximum(add(mul(absval(p.x), (double)0.866025f), mul(p.y, 0.5)), neg(p.y)),
                       cst((double)(1.2f * 0.5f), p.x)));
}
// The knob (SdfFunctionsQueries.cpp:245-267); variant KNOB, KNOB_INNER or KNOB_OUTER.
template <class T>
GPIS_HD T knob(P3<T> p, int variant, const MeanParams &M) {
    const double size = (double)0.8f;
    const double s = (double)(float)(1.0 / size);
    p = P3<T>{mul(p.x, s), mul(p.y, s), mul(p.z, s)};
    if (variant == KNOB_INNER) return mul(sphere(p, 0.75), size);
    const T outer = sphere(p, 1.0);
    const T cutout = sphere(shifted(p, 0.0, -(double)0.5f, -(double)0.5f), (double)0.7f);
    const P3<T> torus = rotated(shifted(p, 0.0, -(double)0.2f, -(double)0.2f), M.rot_m45);
    const T zero = cst(0.0, p.x);
    const T etch = sub(norm(sub(norm(torus.x, torus.z, zero), cst((double)1.0f, p.x)), torus.y, zero),
                       cst((double)0.05f, p.x));
    T d = subtract(cutout, outer, (double)0.1f);
    const T inner = sphere(p, 0.75);
    d = variant == KNOB_OUTER ? maximum(d, neg(inner)) : minimum(d, inner);
    d = maximum(neg(etch), d);
    const P3<T> b = shifted(p, 0.0, (double)0.775f, 0.0);
    T base = smooth_union(cone(rotated(shifted(b, 0.0, (double)0.9f, 0.0), M.rot_m90)),
                          cone(rotated(shifted(b, 0.0, -(double)0.9f, 0.0), M.rot_p90)), (double)0.02f);
    base = mul(maximum(base, cylinder(b, (double)1.1f, (double)0.25f)), (double)0.7f);
    base = maximum(neg(cylinder(b, (double)0.6f, (double)0.3f)), base);
    base = maximum(neg(prism(rotated(shifted(b, 0.0, 0.0, -1.0), M.rot_p90))), base);
    return mul(minimum(subtract(outer, base, (double)0.1f), d), size);
}
// Two spheres of radius 9.5 at y = +-10, folded on |y| (:268-272).
template <class T>
GPIS_HD T two_spheres(const P3<T> &p) {
    return sphere(P3<T>{p.x, sub(absval(p.y), cst(10.0, p.x)), p.z}, 9.5);
}
// A 1.2 x 0.02 x 1.2 slab (:273-280) and the two- and three-slab stacks (:281-291).
template <class T>
GPIS_HD T slab(const P3<T> &p) {
    const T x = sub(absval(p.x), cst(0.6, p.x));
    const T y = sub(absval(p.y), cst(0.01, p.x));
    const T z = sub(absval(p.z), cst(0.6, p.x));
    return maximum(maximum(x, y), z);
}
template <class T>
GPIS_HD T two_slabs(const P3<T> &p) {
    return minimum(slab(shifted(p, -0.18, -0.02, 0.0)), slab(shifted(p, 0.18, 0.02, 0.0)));
}
template <class T>
GPIS_HD T three_slabs(const P3<T> &p) {
    return minimum(two_slabs(p), slab(shifted(p, -0.18, 0.06, 0.0)));
}
// The query geometry dispatch (:293-297): every other function is the knob.
template <class T>
GPIS_HD T geometry(const P3<T> &p, int fn, const MeanParams &M) {
    if (fn == TWO_SLABS) return two_slabs(p);
    if (fn == THREE_SLABS) return three_slabs(p);
    return fn == TWO_SPHERES ? two_spheres(p) : knob(p, KNOB, M);
}

// The upward-open rounded U channel (:323-327), value and gradient.
GPIS_HD Jet rounded_u(double x, double y) {
#pragma clang fp contract(off)
    const double qy = 0.0 < y ? 0.0 : y;
    const double r = sqrt(x * x + qy * qy + 0.0 * 0.0);
    if (r > 0.0) return Jet{1.0 - r, -x / r, -qy / r, -0.0 / r, false, false};
    return Jet{1.0 - r, 0.0, 0.0, 0.0, false, false};
}

// SdfFunctions::evaluateJet at a local point (:444-450).
GPIS_HD Jet evaluate_jet(const MeanParams &M, double x, double y, double z) {
    if (M.fn == ROUNDED_U_GROOVE) return rounded_u(x, y);
    return geometry(P3<Jet>{coord(x, 0), coord(y, 1), coord(z, 2)}, M.fn, M);
}

// SdfFunctions::eval of the functions without queries (SdfFunctions.hpp:64-75): plane is
// p.y, the knob variants are knobGeometry on doubles rounded to float.
GPIS_HD float eval_float(const MeanParams &M, float x, float y, float z) {
    if (M.fn == PLANE) return y;
    return (float)knob(P3<double>{(double)x, (double)y, (double)z}, M.fn, M);
}

// ---------------------------------------------------------------- the mean --
// ProceduralMean::mean for a function without queries (GPFunctions.cpp:282-289): the
// point rounded to float, the float inverse transform, the float API, scale, offset, min.
GPIS_HD double mean_plain(const MeanParams &M, double ax, double ay, double az) {
#pragma clang fp contract(off)
    const float px = (float)ax, py = (float)ay, pz = (float)az;
    const float *m = M.inv;
    const float lx = m[0] * px + m[1] * py + m[2] * pz + m[3];
    const float ly = m[4] * px + m[5] * py + m[6] * pz + m[7];
    const float lz = m[8] * px + m[9] * py + m[10] * pz + m[11];
    float v = eval_float(M, lx, ly, lz);
    v = v * M.scale;
    v = v + M.offset;
    return (double)(M.minv > v ? M.minv : v);
}

// ProceduralMean::evaluateJet (ProceduralMeanSdf.cpp:32-38): the local point W a + t (the
// order of the reference's 3x3 double product: rows 0 and 1 summed left to right, row 2
// as x + (y + z)), then scale, offset and the floor at min; the gradient s W^T g, with
// the scale folded into W's entries first and each row summed left to right.
GPIS_HD double mean_jet(const MeanParams &M, double ax, double ay, double az, double g[3]) {
#pragma clang fp contract(off)
    const double *w = M.w;
    const double lx = w[0] * ax + w[1] * ay + w[2] * az + M.t[0];
    const double ly = w[3] * ax + w[4] * ay + w[5] * az + M.t[1];
    const double lz = w[6] * ax + (w[7] * ay + w[8] * az) + M.t[2];
    const Jet q = evaluate_jet(M, lx, ly, lz);
    const double value = q.v * (double)M.scale + (double)M.offset;
    if (value < (double)M.minv || M.scale == 0.0f) {
        g[0] = g[1] = g[2] = 0.0;
        return value < (double)M.minv ? (double)M.minv : value;
    }
    const double s = (double)M.scale;
    for (int i = 0; i < 3; ++i) g[i] = (s * w[i]) * q.g0 + (s * w[3 + i]) * q.g1 + (s * w[6 + i]) * q.g2;
    return value;
}

// ProceduralMean::mean and dmean_da (GPFunctions.cpp:282-298) at a double point. g may be
// null when only the value is wanted.
GPIS_HD double mean(const MeanParams &M, double ax, double ay, double az, double *g) {
#pragma clang fp contract(off)
    if (M.queries) {
        double gg[3];
        const double v = mean_jet(M, ax, ay, az, gg);
        if (g) { g[0] = gg[0]; g[1] = gg[1]; g[2] = gg[2]; }
        return v;
    }
    const double v0 = mean_plain(M, ax, ay, az);
    if (g) {
        // forward differences with eps = double(0.001f), the other coordinates + 0.0
        const double eps = (double)0.001f;
        const double v1 = mean_plain(M, ax + eps, ay + 0.0, az + 0.0);
        const double v2 = mean_plain(M, ax + 0.0, ay + eps, az + 0.0);
        const double v3 = mean_plain(M, ax + 0.0, ay + 0.0, az + eps);
        g[0] = (v1 - v0) / eps;
        g[1] = (v2 - v0) / eps;
        g[2] = (v3 - v0) / eps;
    }
    return v0;
}

}  // namespace sdf
}  // namespace gpis
