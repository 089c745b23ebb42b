// gpis_procnoise.hpp — the "noise" fields of a proc_nonstationary covariance on the GPU:
// ProceduralNoise (the variance field, GPFunctions.cpp:53-85) and ProceduralNoiseVec (the
// length-scale field, :87-119) over the simplex fbm of SdfFunctions.cpp:84-183.
//
// The ramp types (bottom_top, left_right, front_back, bottom_top_left_right) interpolate
// geometrically between min + 1 and max + 1 along one coordinate; sandstone and rust
// compose three fbm evaluations (2 octaves for the variance, 10 for the length scale).
// Arithmetic follows the reference's types: the field in double, simplex3d and fbm in
// float, random3's sine of the float dot product in double (the reference build resolves
// the unqualified sin to the double one; its value times 4096 rounds to float, whose last
// bits the discontinuous hash then amplifies, so the sine is taken to full double accuracy).
#pragma once

#ifndef GPIS_HOST_CHECK
#include <hip/hip_runtime.h>
#endif

namespace gpis {

// One field: its type (GPIS_NOISE_*) and the JSON min, max with the derived scale and
// offset of each axis (_scale = 1 / (end - start), _offset = -start * _scale).
struct NoiseField {
    int type;
    double mn, mx, scale, offset;
    double mn2, mx2, scale2, offset2;
};

namespace pnoise {

__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
#pragma clang fp contract(off)
    float s = ax * bx;
    s += ay * by;
    s += az * bz;
    return s;
}

// random3 (SdfFunctions.cpp:84-93): discontinuous pseudo-random point of [-0.5, 0.5]^3.
__device__ __forceinline__ void random3(float cx, float cy, float cz, float &rx, float &ry, float &rz) {
#pragma clang fp contract(off)
    const float dd = dot3f(cx, cy, cz, 17.0f, 59.4f, 15.0f);
    float j = (float)(4096.0 * sin((double)dd));
    double t = 512.0 * (double)j;
    const float z = (float)(t - floor(t));
    j = (float)((double)j * 0.125);
    t = 512.0 * (double)j;
    const float x = (float)(t - floor(t));
    j = (float)((double)j * 0.125);
    t = 512.0 * (double)j;
    const float y = (float)(t - floor(t));
    rx = x - 0.5f; ry = y - 0.5f; rz = z - 0.5f;
}

// simplex3d (:110-156): the four surflets of the simplex around p.
__device__ __noinline__ float simplex3d(float px, float py, float pz) {
#pragma clang fp contract(off)
    const float F3 = 0.3333333f, G3 = 0.1666667f;
    const float dp = dot3f(px, py, pz, F3, F3, F3);
    const float sx = floorf(px + dp), sy = floorf(py + dp), sz = floorf(pz + dp);
    const float ds = dot3f(sx, sy, sz, G3, G3, G3);
    const float x0 = (px - sx) + ds, y0 = (py - sy) + ds, z0 = (pz - sz) + ds;
    const float ex = (x0 - y0) < 0.0f ? 0.0f : 1.0f;
    const float ey = (y0 - z0) < 0.0f ? 0.0f : 1.0f;
    const float ez = (z0 - x0) < 0.0f ? 0.0f : 1.0f;
    // i1 = e * (1 - e.zxy), i2 = 1 - e.zxy * (1 - e)
    const float i1x = ex * (1.0f - ez), i1y = ey * (1.0f - ex), i1z = ez * (1.0f - ey);
    const float i2x = 1.0f - ez * (1.0f - ex), i2y = 1.0f - ex * (1.0f - ey), i2z = 1.0f - ey * (1.0f - ez);
    const float g2 = 2.0f * G3, g3 = 3.0f * G3;
    const float x1 = (x0 - i1x) + G3, y1 = (y0 - i1y) + G3, z1 = (z0 - i1z) + G3;
    const float x2 = (x0 - i2x) + g2, y2 = (y0 - i2y) + g2, z2 = (z0 - i2z) + g2;
    const float x3 = (x0 - 1.0f) + g3, y3 = (y0 - 1.0f) + g3, z3 = (z0 - 1.0f) + g3;
    float w[4] = {dot3f(x0, y0, z0, x0, y0, z0), dot3f(x1, y1, z1, x1, y1, z1), dot3f(x2, y2, z2, x2, y2, z2),
                  dot3f(x3, y3, z3, x3, y3, z3)};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float a = 0.6f - w[i];
        w[i] = a < 0.0f ? 0.0f : a;
    }
    float d[4], rx, ry, rz;
    random3(sx, sy, sz, rx, ry, rz);
    d[0] = dot3f(rx, ry, rz, x0, y0, z0);
    random3(sx + i1x, sy + i1y, sz + i1z, rx, ry, rz);
    d[1] = dot3f(rx, ry, rz, x1, y1, z1);
    random3(sx + i2x, sy + i2y, sz + i2z, rx, ry, rz);
    d[2] = dot3f(rx, ry, rz, x2, y2, z2);
    random3(sx + 1.0f, sy + 1.0f, sz + 1.0f, rx, ry, rz);
    d[3] = dot3f(rx, ry, rz, x3, y3, z3);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] *= w[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] *= w[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) d[i] *= w[i];
    float r = d[0] * 52.0f;
    r += d[1] * 52.0f;
    r += d[2] * 52.0f;
    r += d[3] * 52.0f;
    return r;
}

// fbm (:159-183): octaves + 1 simplex layers at frequencies 0.5 * 2.1042^i after the base.
__device__ __forceinline__ double fbm(double ux, double uy, double uz, int octaves) {
#pragma clang fp contract(off)
    const float gain = 0.65f, lacunarity = 2.1042f;
    float frequency = 0.5f, amplitude = gain;
    const float x = (float)(ux * 5.0), y = (float)(uy * 5.0), z = (float)(uz * 5.0);
    float total = simplex3d(x, y, z);
    for (int i = 0; i < octaves; ++i) {
        total += simplex3d(x * frequency, y * frequency, z * frequency) * amplitude;
        frequency *= lacunarity;
        amplitude *= gain;
    }
    total = (float)(((double)total + 2.0) / 4.0);
    return (double)total;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
    const double a = v < lo ? lo : v;
    return a > hi ? hi : a;
}
__device__ __forceinline__ double lerpd(double a, double b, double t) {
#pragma clang fp contract(off)
    return a * (1.0 - t) + b * t;
}
// sqrt(exp(lerp(log(min^2), log(max^2), clamp(c * scale + offset, 0, 1))))
__device__ __forceinline__ double ramp(double mn, double mx, double c, double scale, double offset) {
#pragma clang fp contract(off)
    return sqrt(exp(lerpd(log(mn * mn), log(mx * mx), clampd(c * scale + offset, 0.0, 1.0))));
}

}  // namespace pnoise

// ProceduralNoise::operator() (vec false) or the largest float component of
// ProceduralNoiseVec::operator() (vec true) at p; _const = 1.  Out of line: the fbm types
// are long and only the nonstationary kernels call it.
__device__ __noinline__ double noise_field(const NoiseField &F, bool vec, double px, double py, double pz) {
#pragma clang fp contract(off)
    using namespace pnoise;
    const double cst = 1.0;
    const double mn = F.mn + cst, mx = F.mx + cst;
    double v;
    switch (F.type) {
    case 0: v = ramp(mn, mx, py, F.scale, F.offset) - cst; break;   // bottom_top
    case 1: v = ramp(mn, mx, px, F.scale, F.offset) - cst; break;   // left_right
    case 2: v = ramp(mn, mx, pz, F.scale, F.offset) - cst; break;   // front_back
    case 3: {                                                        // bottom_top_left_right
        const float bt = (float)ramp(mn, mx, py, F.scale, F.offset);
        const float lr = (float)ramp(F.mn2 + cst, F.mx2 + cst, px, F.scale2, F.offset2);
        v = (double)(bt * lr) - cst * cst;
        break;
    }
    case 4: {                                                        // sandstone
        const int oct = vec ? 10 : 2;
        const double sx = px * 0.3, sy = py * 0.3, sz = pz * 0.3;
        const double a = fbm(sx, sy, sz, oct);
        const double b = fbm(sx + a, sy + a, sz + a, oct);
        const double f = fbm(sx + b, sy + b, sz + b, oct);
        if (!vec) {
            v = lerpd(F.mn, F.mx, clampd(sqrt(f * 1.9 * 1.2) - 0.35, 0., 1.));
        } else {
            const float cx = (float)clampd((sqrt(f * 1.9 * 1.2) - 0.35) * 0.2, 0., 1.);
            const float cy = (float)clampd((sqrt(f * 0.7 * 1.2) - 0.35) * 0.2, 0., 1.);
            const float cz = (float)clampd((sqrt(f * 0.25 * 1.2) - 0.35) * 0.2, 0., 1.);
            return (double)fmaxf(fmaxf(cx, cy), cz);
        }
        break;
    }
    default: {                                                       // rust
        const int oct = vec ? 10 : 2;
        const double sx = px * 2.0, sy = py * 2.0, sz = pz * 2.0;
        const double in = fbm(sx * .1, sy * .1, sz * .1, oct) * 0.4;
        const double a = fbm(sx + in, sy + in, sz + in, oct);
        const double c = fbm(sx * 25., sy * 25., sz * 25., oct) * 0.1;
        double x = vec ? a + c : a - c;
        x = clampd((x - 0.4) / (0.6 - 0.4), 0.0, 1.0);
        const double f = x * x * (3.0 - 2.0 * x);
        if (!vec) {
            v = lerpd(F.mn, F.mx, clampd(f, 0., 1.));
        } else {
            const float cx = (float)(0.278 * (1.0 - f) + 1.0 * f);
            const float cy = (float)(0.212 * (1.0 - f) + 1.0 * f);
            const float cz = (float)(0.141 * (1.0 - f) + 1.0 * f);
            return (double)fmaxf(fmaxf(cx, cy), cz);
        }
        break;
    }
    }
    return vec ? (double)(float)v : v;
}

}  // namespace gpis
