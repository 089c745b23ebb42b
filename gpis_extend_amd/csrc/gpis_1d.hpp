// gpis_1d.hpp — the 1D along-ray sparse-convolution flight on gfx950 (Renewal
// model): 1D noise in isotropic ray space, Renewal/Renewal+ conditioning, the
// transverse gradient drawn by Box-Muller in FP64, and next-event estimation
// (neeGrad / neePDF) toward a point light.
//
// Reference behaviour restated (src/core/math/SparseConvolutionNoise.cpp):
//   noise1D / cell1D                       :454-485
//   evaluateNoise1DNormalized              :379-417
//   evaluateGradientNoise1D                :178-243 (single level, no transverse law)
//   conditioning1D                         :576-656
//   neeShared / neeGrad / neePDF           :662-821
// 1D kernels: GPFunctions.cpp:582-605, 845-875.  Box-Muller: Gaussian.cpp:22-34.
#pragma once

#include "gpis_device.hpp"
#include "gpis_exp.hpp"

namespace gpis {

constexpr float kPiF = 3.1415926536f;  // Tungsten's float PI (Angle.hpp:8)

struct Frame3 {
    float nx, ny, nz;  // normal (the ray direction in isotropic space)
    float tx, ty, tz;  // tangent
    float bx, by, bz;  // bitangent
};

// Vec::normalized: multiply by 1/length with length = sqrt(((x*x + y*y) + z*z)).
__device__ __forceinline__ void normalize3(float &x, float &y, float &z) {
#pragma clang fp contract(off)
    float l2 = x * x;
    l2 += y * y;
    l2 += z * z;
    float inv = 1.0f / sqrtf(l2);
    x *= inv; y *= inv; z *= inv;
}
// Vec::dot: sum = a0*b0; sum += a1*b1; sum += a2*b2.
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
#pragma clang fp contract(off)
    float s = ax * bx;
    s += ay * by;
    s += az * bz;
    return s;
}
// TangentFrame(n) (TangentFrame.hpp:74-85), the Duff et al. basis.
__device__ __forceinline__ Frame3 make_frame(float x, float y, float z) {
#pragma clang fp contract(off)
    Frame3 F;
    normalize3(x, y, z);
    F.nx = x; F.ny = y; F.nz = z;
    float sign = copysignf(1.0f, z);
    const float a = -1.0f / (sign + z);
    const float b = x * y * a;
    F.tx = 1.0f + sign * x * x * a; F.ty = sign * b; F.tz = -sign * x;
    F.bx = b; F.by = sign + y * y * a; F.bz = -y;
    normalize3(F.tx, F.ty, F.tz);
    normalize3(F.bx, F.by, F.bz);
    return F;
}

// The mean and its gradient evaluated in double as the reference does
// (GaussianProcess::mean, SphericalMean GPFunctions.hpp:952-996, and the box/plane
// signed distances), the gradient rounded to float.
// PROC = false leaves the procedural mean's out-of-line call out of a kernel (a call
// changes the register allocation of the whole kernel); mean_kind 4 then must not reach it.
template <bool PROC = true>
__device__ __forceinline__ double mean_value_d(const FieldParams &P, float x, float y, float z) {
#pragma clang fp contract(off)
    if (P.mean_kind == 3) return grid_mean_d(P, x, y, z);
    if (PROC && P.mean_kind == 4) return proc_mean_d(P.sdf, (double)x, (double)y, (double)z, nullptr);
    const double p[3] = {(double)x - P.mcd[0], (double)y - P.mcd[1], (double)z - P.mcd[2]};
    if (P.mean_kind == 1) {
        double o2 = 0.0, mx = -1e300;
        for (int i = 0; i < 3; ++i) {
            const double q = fabs(p[i]) - P.mhd[i];
            const double c = q > 0 ? q : 0.0;
            o2 += c * c;
            if (q > mx) mx = q;
        }
        return sqrt(o2) + (mx < 0 ? mx : 0.0);
    }
    if (P.mean_kind == 2) {
        return P.pnd[0] * (double)x + P.pnd[1] * (double)y + P.pnd[2] * (double)z - P.poffd;
    }
    double s = 0.0;
    s += p[0] * p[0];
    s += p[1] * p[1];
    s += p[2] * p[2];
    return sqrt(s) - P.mrd;
}
// The analytic means in FP32 (sphere, box, plane); the tabulated and procedural means keep
// their double evaluation, rounded once.
template <bool PROC = true>
__device__ __forceinline__ float mean_value_f(const FieldParams &P, float x, float y, float z) {
    if (P.mean_kind == 3 || P.mean_kind == 4) return (float)mean_value_d<PROC>(P, x, y, z);
    const float px = x - (float)P.mcd[0], py = y - (float)P.mcd[1], pz = z - (float)P.mcd[2];
    if (P.mean_kind == 1) {
        const float qx = fabsf(px) - (float)P.mhd[0], qy = fabsf(py) - (float)P.mhd[1], qz = fabsf(pz) - (float)P.mhd[2];
        const float cx = fmaxf(qx, 0.f), cy = fmaxf(qy, 0.f), cz = fmaxf(qz, 0.f);
        return sqrtf(cx * cx + cy * cy + cz * cz) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.f);
    }
    if (P.mean_kind == 2)
        return (float)P.pnd[0] * x + (float)P.pnd[1] * y + (float)P.pnd[2] * z - (float)P.poffd;
    return sqrtf(px * px + py * py + pz * pz) - (float)P.mrd;
}
template <bool PROC = true>
__device__ __forceinline__ void mean_grad_d(const FieldParams &P, float x, float y, float z, float &gx, float &gy,
                                            float &gz) {
#pragma clang fp contract(off)
    const double p[3] = {(double)x - P.mcd[0], (double)y - P.mcd[1], (double)z - P.mcd[2]};
    double g[3] = {0.0, 0.0, 0.0};
    if (P.mean_kind == 3) {
        grid_mean_grad_d(P, x, y, z, g[0], g[1], g[2]);
    } else if (PROC && P.mean_kind == 4) {
        proc_mean_d(P.sdf, (double)x, (double)y, (double)z, g);
    } else if (P.mean_kind == 1) {
        double q[3], o2 = 0.0;
        int imax = 0;
        for (int i = 0; i < 3; ++i) {
            q[i] = fabs(p[i]) - P.mhd[i];
            if (q[i] > q[imax]) imax = i;
            if (q[i] > 0) o2 += q[i] * q[i];
        }
        if (o2 > 0) {
            const double inv = 1.0 / sqrt(o2);
            for (int i = 0; i < 3; ++i) g[i] = (q[i] > 0 ? q[i] : 0.0) * inv * (p[i] < 0 ? -1.0 : 1.0);
        } else {
            g[imax] = p[imax] < 0 ? -1.0 : 1.0;
        }
    } else if (P.mean_kind == 2) {
        g[0] = P.pnd[0]; g[1] = P.pnd[1]; g[2] = P.pnd[2];
    } else {
        double s = 0.0;
        s += p[0] * p[0];
        s += p[1] * p[1];
        s += p[2] * p[2];
        const double inv = 1.0 / sqrt(s);
        g[0] = p[0] * inv; g[1] = p[1] * inv; g[2] = p[2] * inv;
    }
    gx = (float)g[0]; gy = (float)g[1]; gz = (float)g[2];
}

// Ray-space scales of one flight: the medium's own, or those of a per-ray SE length
// scale l (SquaredExponentialCovariance::fromJson arithmetic).
__device__ __forceinline__ void ray_scales(const float aniso[3], const float la_med[3], const float ia_med[3],
                                           const float *lscale, unsigned r, float la[3], float ia[3]) {
#pragma clang fp contract(off)
    if (!lscale) {
        for (int i = 0; i < 3; ++i) { la[i] = la_med[i]; ia[i] = ia_med[i]; }
        return;
    }
    const float lc = lscale[r] * sqrtf(2.f) / 2;
    for (int i = 0; i < 3; ++i) {
        la[i] = lc * aniso[i];
        const float inv = 1.0f / la[i];
        ia[i] = isfinite(inv) ? inv : 0.0f;
    }
}

// Per-flight constants of the 1D field: isotropic scaling, ray frame, and the
// coordinate of the conditioning origin.
struct Ray1D {
    float la[3];    // l_aniso = l_conv * aniso, l_conv = l * sqrt(2) / 2 (GPFunctions.cpp:661-668)
    float ia[3];    // 1 / l_aniso (0 where infinite)
    Frame3 F;       // frame of normalize(worldToLocal(dir))
    uint32_t seed;
    int cond;       // conditioning splats active
    float uo;       // origin coordinate in isotropic ray space
    float v1, g1z;  // coeff_1D.value_scale and gradient_scale.z
};

// Coordinate along the ray in isotropic ray space: frame.toLocal(worldToLocal(p)).z
// (SparseConvolutionNoise.cpp:381-384), rounded as the reference.
__device__ __forceinline__ float ray_coord(const Ray1D &R, float px, float py, float pz) {
#pragma clang fp contract(off)
    float ix = R.ia[0] * px, iy = R.ia[1] * py, iz = R.ia[2] * pz;
    return dot3(R.F.nx, R.F.ny, R.F.nz, ix, iy, iz);
}

// Three cells of 1D impulses kept in registers: (t_i with the weight in its sign bit).
template <int NI>
struct Cells1D {
    int base;                 // cell index of the middle cell (i)
    float s[3][NI > 0 ? NI : 1];
};

template <int NI>
__device__ __forceinline__ void gen_cell1d(float *dst, int cell, uint32_t seed) {
    Pcg g;
    g.set((uint64_t)(uint32_t)(xxh2((uint32_t)cell, seed) + 1u));  // cell1D (:466)
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        float t = u01(g.next());
        float w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
        dst[k] = copysignf(t, w);
    }
}

// Make the cache hold cells i-1, i, i+1, shifting when the ray advanced one cell.
template <int NI>
__device__ __forceinline__ void cells_for(Cells1D<NI> &C, int i, uint32_t seed, bool valid) {
    if (valid && i == C.base) return;
    if (valid && i == C.base + 1) {
#pragma unroll
        for (int k = 0; k < NI; ++k) { C.s[0][k] = C.s[1][k]; C.s[1][k] = C.s[2][k]; }
        gen_cell1d<NI>(C.s[2], i + 1, seed);
    } else {
        gen_cell1d<NI>(C.s[0], i - 1, seed);
        gen_cell1d<NI>(C.s[1], i, seed);
        gen_cell1d<NI>(C.s[2], i + 1, seed);
    }
    C.base = i;
}

// exp(x) for a float argument, rounded to float from double as the reference's
// kernels compute it ((float)exp((double)x), GPFunctions.cpp:845-875).
__device__ __forceinline__ float exp_dr(float x) { return (float)exp((double)x); }

// exp(-x^2 / 2) in FP32 by the hardware exponential: exp2(-x^2 log2(e) / 2).
constexpr float kNegHalfLog2e = -0.72134752044448170f;
__device__ __forceinline__ float gauss_f32(float x) { return hw_exp2(kNegHalfLog2e * (x * x)); }

// noise1D (:454-485) at coordinate u with kernel radius R = 3: the raw sums of w*h and
// w*dh/du.  FAST = false: bit-identical to the reference's arithmetic (per-cell partial
// sums added cell by cell, splat products rounded as written, exponentials rounded from
// double).  FAST = true: the same support decisions and summation order, the splat by
// the FP32 hardware exponential (a few ulp).
template <int NI, bool FAST = false>
__device__ __forceinline__ void noise1d_raw(Cells1D<NI> &C, bool &cache_valid, float u, uint32_t seed,
                                            float &v, float &dv) {
#pragma clang fp contract(off)
    const float R = 3.0f;
    float tg = u / R;
    float fl = floorf(tg);
    float fr = tg - fl;
    int i = (int)fl;
    cells_for<NI>(C, i, seed, cache_valid);
    cache_valid = true;
    float sv = 0.f, sd = 0.f;
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
        const float t = fr - (float)(dx - 1);
        const float pq = R * t;
        float cv = 0.f, cd = 0.f;
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const float ts = C.s[dx][k];
            const float ti = fabsf(ts);
            const float to = t - ti;
            if (to * to < 1.0f) {
                const float pc = R * ti;
                const float ab = pq - pc;                // splattingKernel1D: pQuery - pCenter
                const float e = FAST ? gauss_f32(ab) : (float)exp_splat_d((double)(-(ab * ab / 2.0f)));
                const float g = -ab * e;                 // (-2 ab / 2) * e, exact in double then rounded
                const float w = copysignf(1.0f, ts);
                cv += w * e;
                cd += w * g;
            }
        }
        sv += cv;
        sd += cd;
    }
    v = sv;
    dv = sd;
}

// Covariance splats of the 1D conditioning (covarianceKernel1D / 1DGrad with the
// stationary kernel of length sqrt(2), GPFunctions.cpp:590-605, 845-875), rounded as
// the reference: value v1 * kappa + g1z * kappa', derivative v1 * kappa' + g1z * kappa''.
template <bool FAST = false>
__device__ __forceinline__ void cond_splat1d(const Ray1D &R, float u, float &v, float &dv) {
#pragma clang fp contract(off)
    const float d = u - R.uo;
    if (FAST) {
        // kappa = exp(-d^2 / 4), kappa' = -d/2 kappa, kappa'' = (d^2/4 - 1/2) kappa in FP32
        const float k = hw_exp2(-0.36067376022224085f * (d * d));
        const float h = -0.5f * d;
        const float k1 = h * k, k2 = (h * h - 0.5f) * k;
        v += R.v1 * k + R.g1z * k1;
        dv += R.v1 * k1 + R.g1z * k2;
        return;
    }
    const float den2 = 3.9999998f;                     // (float)(2 * sqr((float)(1 * sqrt(2))))
    const float k = exp_dr(-(d * d / den2));
    const float k1 = (float)((-2.0 * (double)d) / (double)den2 * (double)k);
    const float g1 = (float)(-2.0 * (double)d / 4.0);
    const float k2 = (float)(((double)(g1 * g1) - 0.5) * (double)k);
    const float vs = R.v1 * k, gs = R.g1z * k1;
    const float vd = R.v1 * k1, gd = R.g1z * k2;
    v += vs + gs;
    dv += vd + gd;
}

// Box-Muller pair m (0 or 1) after set_state(normal_seed), in FP64 (Gaussian.cpp:22-34).
__device__ __forceinline__ void normal_pair(uint32_t normal_seed, int m, double &x, double &y) {
    Pcg g;
    g.set((uint64_t)normal_seed);
    double u1 = u01(g.next()), u2 = u01(g.next());
    if (m == 1) { u1 = u01(g.next()); u2 = u01(g.next()); }
    double r = sqrt(-2.0 * log(1.0 - u1));
    double a = (double)(2.0f * kPiF) * u2;
    x = r * cos(a);
    y = r * sin(a);
}
// The same pair divided by sqrt(2.f) as evaluateGradientNoise1D does (:204-222).
__device__ __forceinline__ void transverse_pair(uint32_t normal_seed, int m, double &x, double &y) {
    normal_pair(normal_seed, m, x, y);
    const double s2 = (double)sqrtf(2.0f);
    x = x / s2;
    y = y / s2;
}

// seTransverseGradientMoments (SparseConvolutionNoise.cpp:78-99) for a 1D SE Renewal+
// medium with 1D_gradient_correlationXY: the transverse gradient at p is Gaussian with
// mean rho * (observed transverse gradient at the origin) and variance
// -expm1(-q2) / 2, q2 = |T (p - o)|^2 / 2; without a gradient observation mean 0 and
// variance 1/2.  obs_x, obs_y: frame.toLocal(localToWorld^T g_obs) of the conditioning.
__device__ __forceinline__ void transverse_moments(const Ray1D &R, bool observed, float obs_x, float obs_y, float ox,
                                                   float oy, float oz, float px, float py, float pz, double &mx,
                                                   double &my, double &var) {
#pragma clang fp contract(off)
    mx = 0.0; my = 0.0; var = 0.5;
    if (!observed) return;
    const float ex = R.ia[0] * (px - ox), ey = R.ia[1] * (py - oy), ez = R.ia[2] * (pz - oz);
    double l2 = (double)ex * (double)ex;
    l2 += (double)ey * (double)ey;
    l2 += (double)ez * (double)ez;
    const double q2 = 0.5 * l2;
    const double rho = exp(-0.5 * q2);
    var = -0.5 * expm1(-q2);
    mx = rho * (double)obs_x;
    my = rho * (double)obs_y;
}

// World gradient of the noise from (transverse x, y; longitudinal dv) in isotropic
// ray space: frame.toGlobal then localToWorld^T (evaluateGradientNoise1D :240-242).
__device__ __forceinline__ void ray_grad_to_world(const Ray1D &R, float gx, float gy, float gz, float &wx, float &wy,
                                                  float &wz) {
#pragma clang fp contract(off)
    float x = R.F.tx * gx + R.F.bx * gy;
    x += R.F.nx * gz;
    float y = R.F.ty * gx + R.F.by * gy;
    y += R.F.ny * gz;
    float z = R.F.tz * gx + R.F.bz * gy;
    z += R.F.nz * gz;
    wx = R.ia[0] * x; wy = R.ia[1] * y; wz = R.ia[2] * z;
}

}  // namespace gpis
