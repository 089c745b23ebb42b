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

// -log2(e) / 2: exp(-x^2 / 2) = exp2(kNegHalfLog2e x^2) on the hardware exponential.
constexpr float kNegHalfLog2e = -0.72134752044448170f;

// The sums of noise1D (:454-485) over the three cells around the query, given the
// impulses of cells i-1, i, i+1 (s[dx][k], t_k with the weight in its sign bit) and the
// query's fraction fr of cell i: the raw sums of w*h and w*dh/du with kernel radius R = 3.
// FAST = false: bit-identical to the reference's arithmetic (per-cell partial sums added
// cell by cell, splat products rounded as written, exponentials rounded from double).
// FAST = true: the same support decisions and summation order, the splat by the FP32
// hardware exponential (a few ulp).
template <int NI, bool FAST = false>
__device__ __forceinline__ void noise1d_sum(const float (&s)[3][NI > 0 ? NI : 1], float fr, float &v, float &dv) {
#pragma clang fp contract(off)
    const float R = 3.0f;
    float sv = 0.f, sd = 0.f;
    if (FAST) {
        // the splat exp(-ab^2 / 2) with ab = R to, from the to^2 of the support test:
        // exp2(-log2(e) R^2 / 2 * to^2); the derivative sum w e (-ab) as -R * sum w e to
        const float kq = kNegHalfLog2e * (R * R);
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const float t = fr - (float)(dx - 1);
            float cv = 0.f, cd = 0.f;
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                const float ts = s[dx][k];
                const float to = t - fabsf(ts);
                const float q = to * to;
                const float e = hw_exp2(kq * q);
                const float we = q < 1.0f ? copysignf(e, ts) : 0.0f;
                cv += we;
                cd += we * to;
            }
            sv += cv;
            sd += cd;
        }
        v = sv;
        dv = -R * sd;
        return;
    }
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
        const float t = fr - (float)(dx - 1);
        const float pq = R * t;
        float cv = 0.f, cd = 0.f;
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const float ts = s[dx][k];
            const float ti = fabsf(ts);
            const float to = t - ti;
            if (to * to < 1.0f) {
                const float pc = R * ti;
                const float ab = pq - pc;                // splattingKernel1D: pQuery - pCenter
                const float e = (float)exp_splat_d((double)(-(ab * ab / 2.0f)));
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

// noise1D at coordinate u with the lane's register cache of three cells.
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
    noise1d_sum<NI, FAST>(C.s, fr, v, dv);
}

// ------------------------------------------------------------------------------------------
// The cells of the 1D flight kernel in LDS, generated by the whole wave.
//
// A lane's three cells i-1, i, i+1 live in a ring of three LDS slots (cell c in slot c mod 3),
// so advancing one cell replaces one slot and stepping back (a secant point) keeps two.  A
// cell is NI impulses from cell1D's PCG stream (set_state(xxhash32(c, seed) + 1), then per
// impulse t and the weight's draw).  Draw j of a stream is the output of the state after
// j + 2 steps (set_state discards two draws), and PCG's state after m steps is a^m s + c_m
// (c_m = a^(m-1) + ... + 1), so any impulse of any cell is two multiply-adds of the seeded
// state: the missing cells of all lanes of the wave are cut into (cell, impulse) items, and
// every active lane makes one item per round.  The values are those of the sequential
// generator to the bit; a lane no longer runs a whole cell's chain while the wave waits.
//
// LDS layout per wave: [slot 3][impulse quad NI/4][lane 64] float4, read as two 16-byte
// loads per cell with consecutive lanes at consecutive addresses (no bank conflicts).
constexpr uint64_t kPcgMult = 6364136223846793005ULL;
constexpr int kJumpSteps = 2 * 8 + 2;  // states 0 .. 2 NI + 1 of a cell's stream (NI = 8)
struct PcgJump {
    uint64_t a[kJumpSteps], c[kJumpSteps];
};
constexpr PcgJump make_pcg_jump() {
    PcgJump J{};
    uint64_t a = 1, c = 0;
    for (int m = 0; m < kJumpSteps; ++m) {
        J.a[m] = a;
        J.c[m] = c;
        c = c * kPcgMult + 1ULL;
        a = a * kPcgMult;
    }
    return J;
}
__constant__ PcgJump kPcgJumpTable = make_pcg_jump();
// PCG32's output of the state `old` (the XSH-RR permutation of Pcg::next).
__device__ __forceinline__ uint32_t pcg_output(uint64_t old) {
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
// The ring slot of cell c.
__device__ __forceinline__ int ring_slot(int c) {
    const int m = c % 3;
    return m < 0 ? m + 3 : m;
}

template <int NI>
struct Cells1DShared {
    static constexpr int kQuads = NI / 4;
    static constexpr int kMaxJobs = 3 * 64;
    float4 *cells;       // this wave's [3][kQuads][64] float4
    uint32_t *job_st;    // this wave's seeded states of the cells being made
    uint32_t *job_dst;   // their owner lane | slot << 6
    int base;            // the lane's middle cell
    bool valid;
};

// Make the lane's slots hold cells i-1, i, i+1 (all active lanes of the wave call this
// together; `need` lanes with nothing missing still take part in the rounds).
template <int NI>
__device__ __forceinline__ void cells_shared_for(Cells1DShared<NI> &C, int i, uint32_t seed) {
    static_assert(2 * NI + 2 <= kJumpSteps, "jump table covers a cell's stream");
    int miss[3], nj = 0;
    const int d = i - C.base;
    if (!C.valid || d > 2 || d < -2) {
        miss[0] = i - 1; miss[1] = i; miss[2] = i + 1; nj = 3;
    } else if (d == 1) {
        miss[0] = i + 1; nj = 1;
    } else if (d == 2) {
        miss[0] = i; miss[1] = i + 1; nj = 2;
    } else if (d == -1) {
        miss[0] = i - 1; nj = 1;
    } else if (d == -2) {
        miss[0] = i - 1; miss[1] = i; nj = 2;
    }
    C.base = i;
    C.valid = true;
    const uint64_t b1 = wave_ballot(nj >= 1);
    if (b1 == 0) return;  // wave-uniform
    const uint64_t b2 = wave_ballot(nj >= 2), b3 = wave_ballot(nj >= 3);
    const uint64_t act = wave_ballot(true);
    const unsigned lane = lane64();
    const uint64_t lt = (1ull << lane) - 1ull;
    const unsigned first = (unsigned)(popc64(b1 & lt) + popc64(b2 & lt) + popc64(b3 & lt));
    const unsigned jobs = (unsigned)(popc64(b1) + popc64(b2) + popc64(b3));
    for (int r = 0; r < 3; ++r) {
        if (r < nj) {
            C.job_st[first + r] = xxh2((uint32_t)miss[r], seed) + 1u;  // cell1D (:466)
            C.job_dst[first + r] = lane | ((unsigned)ring_slot(miss[r]) << 6);
        }
    }
    wave_sync();
    // items (job j, impulse k) = (item / NI, item % NI), taken by the active lanes in turn
    const unsigned na = (unsigned)popc64(act), rank = (unsigned)popc64(act & lt);
    const unsigned items = jobs * NI;
    for (unsigned it = rank; it < items; it += na) {
        const unsigned j = it / NI, k = it % NI;
        const uint64_t at = kPcgJumpTable.a[2 * k + 2], ct = kPcgJumpTable.c[2 * k + 2];
        const uint64_t aw = kPcgJumpTable.a[2 * k + 3], cw = kPcgJumpTable.c[2 * k + 3];
        const uint64_t st = (uint64_t)C.job_st[j];
        const uint32_t dst = C.job_dst[j];
        const float t = u01(pcg_output(at * st + ct));
        const float w = u01(pcg_output(aw * st + cw)) < 0.5f ? -1.0f : 1.0f;
        const unsigned owner = dst & 63u, slot = dst >> 6;
        float *q = reinterpret_cast<float *>(&C.cells[(slot * Cells1DShared<NI>::kQuads + k / 4) * 64 + owner]);
        q[k & 3] = copysignf(t, w);
    }
    wave_sync();
}

// noise1D at coordinate u with the cells in LDS.
template <int NI, bool FAST = false>
__device__ __forceinline__ void noise1d_shared(Cells1DShared<NI> &C, float u, uint32_t seed, float &v, float &dv) {
#pragma clang fp contract(off)
    const float R = 3.0f;
    float tg = u / R;
    float fl = floorf(tg);
    float fr = tg - fl;
    int i = (int)fl;
    cells_shared_for<NI>(C, i, seed);
    const unsigned lane = lane64();
    float s[3][NI];
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
        const int slot = ring_slot(i - 1 + dx);
#pragma unroll
        for (int qd = 0; qd < NI / 4; ++qd) {
            const float4 q = C.cells[(slot * (NI / 4) + qd) * 64 + lane];
            s[dx][4 * qd] = q.x; s[dx][4 * qd + 1] = q.y; s[dx][4 * qd + 2] = q.z; s[dx][4 * qd + 3] = q.w;
        }
    }
    noise1d_sum<NI, FAST>(s, fr, v, dv);
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
