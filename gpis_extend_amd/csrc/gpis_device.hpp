// gpis_device.hpp — device-side building blocks of the gfx950 GPIS flight engine.
//
// Integer parts (hash, PCG, unit float, Bernoulli, next3D order) are bit-exact
// with the reference; the noise arithmetic is FP32.
#pragma once

#include <stdint.h>

#ifdef GPIS_HOST_CHECK
// Host build of the field code for the CPU checks under tests/ (one lane per wave, or one
// thread per lane): the few device intrinsics the field evaluation uses, restated.
#include <math.h>
#include <pthread.h>
#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __noinline__
#define __constant__
#define __expf expf
#define __logf logf
#define __powf powf
static inline float rsqrtf(float x) { return 1.0f / sqrtf(x); }
static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }
#else
#include <hip/hip_runtime.h>
#endif

#include "gpis_lists.hpp"
#include "gpis_procnoise.hpp"
#include "gpis_sdf.hpp"

namespace gpis {

// exp2 on the transcendental unit (v_exp_f32), and "some lane of the wave" of a predicate.
// wave_ballot: the lanes of the wave whose predicate holds; lane64: this lane's index;
// popc64: bits set; wave_sync: order this wave's LDS writes before its later reads.
#ifdef GPIS_HOST_CHECK
static inline float hw_exp2(float x) { return exp2f(x); }
#define GPIS_CHECK_INDEX(i, n) gpis_check_index((size_t)(i), (size_t)(n), __FILE__, __LINE__)
// A wave of several lanes for the checks of the wave-cooperative code: one host thread per
// lane, the lanes meeting at a barrier in every ballot and wave_sync (their code calls these
// in the same sequence on every lane).  Without one (host_wave null) each thread is a wave
// of one lane.
struct HostWave {
    pthread_barrier_t bar;
    uint64_t vote;
    unsigned first;  // a lane of the wave, which clears the vote
};
inline HostWave *host_wave = nullptr;
inline thread_local unsigned host_lane = 0;
static inline bool wave_any(bool p) { return p; }
static inline uint64_t wave_ballot(bool p) {
    if (!host_wave) return p ? 1u : 0u;
    pthread_barrier_wait(&host_wave->bar);
    if (host_lane == host_wave->first) __atomic_store_n(&host_wave->vote, 0ull, __ATOMIC_SEQ_CST);
    pthread_barrier_wait(&host_wave->bar);
    if (p) __atomic_fetch_or(&host_wave->vote, 1ull << host_lane, __ATOMIC_SEQ_CST);
    pthread_barrier_wait(&host_wave->bar);
    return __atomic_load_n(&host_wave->vote, __ATOMIC_SEQ_CST);
}
static inline unsigned lane64() { return host_wave ? host_lane : 0u; }
static inline int popc64(uint64_t m) { return __builtin_popcountll(m); }
static inline void wave_sync() {
    if (host_wave) pthread_barrier_wait(&host_wave->bar);
}
#else
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ unsigned lane64() { return __lane_id(); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
#define GPIS_CHECK_INDEX(i, n) ((void)0)
#endif

// ---------------------------------------------------------------- leaves --
// xxhash32 on 1, 2 and 4 words: src/core/math/MathUtil.hpp:179-224.
#define GPIS_P2 2246822519U
#define GPIS_P3 3266489917U
#define GPIS_P4 668265263U
#define GPIS_P5 374761393U
__host__ __device__ __forceinline__ uint32_t rotl17(uint32_t h) { return (h << 17) | (h >> 15); }
__host__ __device__ __forceinline__ uint32_t xx_final(uint32_t h) {
    h = GPIS_P2 * (h ^ (h >> 15));
    h = GPIS_P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
__host__ __device__ __forceinline__ uint32_t xxh1(uint32_t p) { return xx_final(GPIS_P4 * rotl17(p + GPIS_P5)); }
__host__ __device__ __forceinline__ uint32_t xxh2(uint32_t x, uint32_t y) {
    return xx_final(GPIS_P4 * rotl17(y + GPIS_P5 + x * GPIS_P3));
}
__host__ __device__ __forceinline__ uint32_t xxh4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    uint32_t h = GPIS_P4 * rotl17(w + GPIS_P5 + x * GPIS_P3);
    h = GPIS_P4 * rotl17(h + y * GPIS_P3);
    h = GPIS_P4 * rotl17(h + z * GPIS_P3);
    return xx_final(h);
}

// PCG32 XSH-RR with increment 1 (UniformSampler.hpp:41-48).  64-bit state
// arithmetic is done with the 32x32->64 mad the ISA provides.
struct Pcg {
    uint64_t s;
    __host__ __device__ __forceinline__ uint32_t next() {
        uint64_t old = s;
        s = old * 6364136223846793005ULL + 1ULL;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // set_state(h): state = h, then next2D() discards two draws (UniformSampler.hpp:71-75).
    __host__ __device__ __forceinline__ void set(uint64_t st) { s = st; next(); next(); }
};
// BitManip::normalizedUint (BitManip.hpp:47-50).
__host__ __device__ __forceinline__ float u01(uint32_t i) {
    union { uint32_t u; float f; } c;
    c.u = (i >> 9u) | 0x3F800000u;
    return c.f - 1.0f;
}

// ------------------------------------------------------------ bounces --
// sin and cos of 2 pi u for u in [0, 1), in float operations that round the same on the
// host and the device (no library calls, no contraction): the quadrant and octant of u by
// exact reductions, then Taylor polynomials of sin and cos on [0, pi/4] (truncation error
// below 2e-9, Horner form).  Used where a host restatement must reproduce the device bit
// for bit (the diffuse bounce of the 3D paths).
__host__ __device__ __forceinline__ void sincos_turn(float u, float &s, float &c) {
#pragma clang fp contract(off)
    const float x4 = u * 4.0f;                  // exact
    const float qf = floorf(x4);
    const int q = (int)qf & 3;
    const float r = x4 - qf;                    // exact, in [0, 1): angle r * pi / 2 in the quadrant
    const bool swap = r > 0.5f;
    const float m = swap ? 1.0f - r : r;        // exact, in [0, 0.5]
    const float h = m * 1.57079632679489662f;   // angle in [0, pi / 4]
    const float h2 = h * h;
    float ps = -2.50521083854417188e-8f;        // -1/11!
    ps = ps * h2 + 2.75573192239858907e-6f;     //  1/9!
    ps = ps * h2 - 1.98412698412698413e-4f;     // -1/7!
    ps = ps * h2 + 8.33333333333333333e-3f;     //  1/5!
    ps = ps * h2 - 1.66666666666666667e-1f;     // -1/3!
    float sn = ps * h2;
    sn = sn * h;
    sn = sn + h;
    float pc = 2.08767569878680990e-9f;         //  1/12!
    pc = pc * h2 - 2.75573192239858907e-7f;     // -1/10!
    pc = pc * h2 + 2.48015873015873016e-5f;     //  1/8!
    pc = pc * h2 - 1.38888888888888889e-3f;     // -1/6!
    pc = pc * h2 + 4.16666666666666667e-2f;     //  1/4!
    pc = pc * h2 - 0.5f;
    float cs = pc * h2;
    cs = cs + 1.0f;
    const float S = swap ? cs : sn, C = swap ? sn : cs;  // sin, cos of r * pi / 2
    s = q == 0 ? S : (q == 1 ? C : (q == 2 ? -S : -C));
    c = q == 0 ? C : (q == 1 ? -S : (q == 2 ? -C : S));
}

// The direction of a cosine-weighted diffuse bounce about the unit normal n (TangentFrame's
// basis, TangentFrame.hpp:74-85) and the new flight's jitter: PCG32 seeded by
// xxhash32(ray, bounce, bounce_seed, 0) + 1 gives u1, u2 and the jitter; local direction
// (sqrt(u1) cos 2 pi u2, sqrt(u1) sin 2 pi u2, sqrt(1 - u1)).  Correctly rounded square roots
// and divisions only, so the host's restatement gives the same bits.
__host__ __device__ __forceinline__ void diffuse_direction(uint32_t ray, uint32_t bounce, uint32_t bounce_seed,
                                                            float nx, float ny, float nz, float &dx, float &dy,
                                                            float &dz, float &jit) {
#pragma clang fp contract(off)
    Pcg g;
    g.set((uint64_t)(uint32_t)(xxh4(ray, bounce, bounce_seed, 0u) + 1u));
    const float u1 = u01(g.next()), u2 = u01(g.next());
    jit = u01(g.next());
    const float rr = sqrtf(u1);
    float sn, cs;
    sincos_turn(u2, sn, cs);
    const float lx = rr * cs, ly = rr * sn;
    const float lz = sqrtf(fmaxf(0.f, 1.f - u1));
    const float sgn = copysignf(1.0f, nz);
    const float aa = -1.0f / (sgn + nz);
    const float bb = (nx * ny) * aa;
    const float tx = 1.0f + ((sgn * nx) * nx) * aa, ty = sgn * bb, tz = -sgn * nx;
    const float bx = bb, by = sgn + (ny * ny) * aa, bz = -ny;
    float x = tx * lx + bx * ly;
    x = x + nx * lz;
    float y = ty * lx + by * ly;
    y = y + ny * lz;
    float z = tz * lx + bz * ly;
    z = z + nz * lz;
    float l2 = x * x + y * y;
    l2 = l2 + z * z;
    const float inv = 1.0f / sqrtf(l2);
    dx = x * inv; dy = y * inv; dz = z * inv;
}

// The unit normal of a hit from its gradient, as the bounce uses it.
__host__ __device__ __forceinline__ void unit_normal(float gx, float gy, float gz, float &nx, float &ny, float &nz) {
#pragma clang fp contract(off)
    float l2 = gx * gx + gy * gy;
    l2 = l2 + gz * gz;
    const float inv = 1.0f / sqrtf(l2);
    nx = gx * inv; ny = gy * inv; nz = gz * inv;
}

// ------------------------------------------------------------ parameters --
struct FieldParams {
    // covariance
    int kind;            // 0 SE, 1 procedural nonstationary
    int multires;        // _multiResolutionGrid
    float sigma;         // covariance sigma
    float amp_const;     // amplitude when it is spatially constant
    float l_aniso[3], l_aniso_inv[3];
    float aniso[3];  // anisotropy factors of the SE length scale (1D rays rescale by their own l)
    // isotropic ray space of 3D media (isotropic_3D_sampling): kernel radius, exponent
    // factor (-log2 e * R^2 / 2), gradient factor (-R) and 1/sqrt(variance) of the
    // identity-space kernel
    int iso3d;
    float iso_R, iso_c0, iso_g, iso_nf_inv;
    float length_scale;  // the medium's SE lengthScale
    float l_max;         // max(l_aniso) (splattingKernelRadius)
    float kernel_scale;  // _kernelScale (3)
    int ls_field; float ls_params[4]; float ls_max; float ns_var;
    // "noise" fields (gpis_procnoise.hpp): var_field 1 makes the amplitude var(p) * sigma;
    // ls_field 2 takes the lateral scale from ls_noise
    int var_field; NoiseField var_noise, ls_noise;
    int N; float Nf;
    int iso;                  // isotropic kernel: exponent = c * r2
    // full anisotropy matrix (useAnisoMtx; stationary SE in world space, NI = -1 kernels):
    // with M = cinv / 2 (cinv = W^T W, W the world-to-local matrix) the splat in grid units
    // t is exp2(fc . (tx^2, ty^2, tz^2, tx ty, tx tz, ty tz)), fc = -log2 e R^2 (M00, M11,
    // M22, 2 M01, 2 M02, 2 M12); fg = -2 R M takes the moments sum w e t to the world
    // gradient; fk holds cinv / 4 (the covariance splat of the conditioning, 00 11 22 01 02
    // 12) and fhinv the inverse of -2 cinv / 4 (sparseConvNoiseOneOverSecondDerivative)
    int full;
    float fc[6], fg[9], fk[6];
    double fhinv[9];
    float w2l[9], l2w[9];  // W and L, row-major (the isotropic ray space's transforms)
    int seed_add_single;      // floor(log(kss)/log(2.5)) for the single-level grid
    int level_seed_add[65];   // the same for multi-resolution level L = -32 .. 32, indexed L + 32
    // single-level world-space constants (stationary SE)
    float R, invR;
    float c2[3];         // R^2 * m_i : exponent factor per axis in grid units
    float gfac[3];       // -2 R m_i : world-gradient factor per axis
    float nf_inv;        // 1 / sqrt(variance)
    // mean
    int mean_kind; float mc[3]; float mr; float mh[3]; float pn[3]; float poff;
    double mcd[3], mhd[3], pnd[3], poffd, mrd;  // the same mean parameters in double (exact 1D path)
    // tabulated mean (mean_kind 3): dense voxel grid, x fastest
    const float *grid;
    int gorig[3], gdim[3];
    float w2i[12], gcmin[3], gcmax[3], goff, gscale;
    int gvolume;                                 // density -> -log(max(1e-4, density))
    // procedural SDF mean (mean_kind 4): its parameters in device memory
    const sdf::MeanParams *sdf;
    // medium
    uint32_t global_seed; int single; int ctxt; int activate; int one_d; int scheme1d; int corr_xy;
    float step_size; float min_step_f; int max_bounce;
    // precomputed impulse table (single realization)
    // packed impulse table: per cell 3N floats (x with the weight in its sign bit, y, z)
    const float *table; int tlo[3]; int tdim[3]; uint32_t tseed; int has_table;
    uint64_t table_n;  // floats in the table
    // support lists of the table's interior cells (gpis_lists.hpp): fixed-point impulses
    // (4 int32 each), list entries (byte offsets into lfix, groups of four), and per
    // sub-cell (first group, number of groups); lsub sub-cells per cell edge
    const int32_t *lfix; const uint32_t *lent; const uint32_t *lrange; int lsub; int has_lists;
    uint32_t lfix_bytes, lent_n, lrange_n;  // their sizes (bytes, entries, sub-cells)
    int lmargin;        // cells whose lists exist: those at least lmargin cells inside the table
    float lpair2;       // lists serve a second query within sqrt(lpair2) grid units (0: one query)
    int pair_cells;     // two points of a flight in one grid cell share a pass over its block's cells
    int coop_cells;     // the flight kernels make generated cells wave-cooperatively (cells_coop)
};

struct Eval4 { float v, gx, gy, gz; };

// ------------------------------------------------------------------ means --
// SphericalMean (GPFunctions.hpp:952-996); box SDF and plane are defined in code.
// Trilinear density of the tabulated mean's voxel grid at world point p, in the order
// of the reference's dense mean grid (RaySegmentBatch.cu:58-90).
__device__ __forceinline__ float grid_voxel(const FieldParams &P, int x, int y, int z) {
    x = max(0, min(x - P.gorig[0], P.gdim[0] - 1));
    y = max(0, min(y - P.gorig[1], P.gdim[1] - 1));
    z = max(0, min(z - P.gorig[2], P.gdim[2] - 1));
    return P.grid[((size_t)z * P.gdim[1] + y) * P.gdim[0] + x];
}
__device__ __forceinline__ float grid_density(const FieldParams &P, float px, float py, float pz) {
#pragma clang fp contract(off)
    const float *m = P.w2i;
    float qx = m[0] * px + m[1] * py + m[2] * pz + m[3];
    float qy = m[4] * px + m[5] * py + m[6] * pz + m[7];
    float qz = m[8] * px + m[9] * py + m[10] * pz + m[11];
    qx = fminf(fmaxf(qx, P.gcmin[0]), P.gcmax[0]);
    qy = fminf(fmaxf(qy, P.gcmin[1]), P.gcmax[1]);
    qz = fminf(fmaxf(qz, P.gcmin[2]), P.gcmax[2]);
    const int x = (int)floorf(qx), y = (int)floorf(qy), z = (int)floorf(qz);
    const float fx = qx - (float)x, fy = qy - (float)y, fz = qz - (float)z;
    float layer[2];
    for (int dz = 0; dz < 2; ++dz) {
        float row[2];
        for (int dy = 0; dy < 2; ++dy) {
            const float a = grid_voxel(P, x, y + dy, z + dz);
            const float b = grid_voxel(P, x + 1, y + dy, z + dz);
            row[dy] = a + fx * (b - a);
        }
        layer[dz] = row[0] + fy * (row[1] - row[0]);
    }
    return layer[0] + fz * (layer[1] - layer[0]);
}

// TabulatedMean::mean (GPFunctions.cpp:165-172): float lookup point, double arithmetic.
__device__ __forceinline__ double grid_mean_d(const FieldParams &P, float x, float y, float z) {
#pragma clang fp contract(off)
    double res = (double)grid_density(P, x, y, z);
    if (P.gvolume) res = -log(fmax(0.0001, res));
    return (res + (double)P.goff) * (double)P.gscale;
}
// TabulatedMean::dmean_da (:179-190): forward differences with eps = 1e-3, the shifted
// point formed in double and looked up in float.
__device__ __forceinline__ double grid_mean_grad_d(const FieldParams &P, float x, float y, float z, double &gx,
                                                   double &gy, double &gz) {
#pragma clang fp contract(off)
    const double eps = 0.001;
    double c = 0.0;
    // one lookup per iteration (the centre, then +x, +y, +z) keeps the inlined code small
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        const float qx = k == 1 ? (float)((double)x + eps) : x;
        const float qy = k == 2 ? (float)((double)y + eps) : y;
        const float qz = k == 3 ? (float)((double)z + eps) : z;
        const double m = grid_mean_d(P, qx, qy, qz);
        if (k == 0) c = m;
        const double d = (m - c) / eps;
        if (k == 1) gx = d;
        if (k == 2) gy = d;
        if (k == 3) gz = d;
    }
    return c;
}

// Tabulated mean (double, as TabulatedMean::mean returns it) and its gradient rounded to float.
__device__ __forceinline__ double grid_mean_value(const FieldParams &P, float x, float y, float z, float &gx,
                                                  float &gy, float &gz) {
    double dx, dy, dz;
    const double c = grid_mean_grad_d(P, x, y, z, dx, dy, dz);
    gx = (float)dx; gy = (float)dy; gz = (float)dz;
    return c;
}

// The analytic means (sphere, box, plane) and their gradients.
__device__ __forceinline__ float analytic_mean_value(const FieldParams &P, float x, float y, float z, float &gx,
                                                     float &gy, float &gz) {
    if (P.mean_kind == 1) {
        float px = x - P.mc[0], py = y - P.mc[1], pz = z - P.mc[2];
        float qx = fabsf(px) - P.mh[0], qy = fabsf(py) - P.mh[1], qz = fabsf(pz) - P.mh[2];
        float ox = fmaxf(qx, 0.f), oy = fmaxf(qy, 0.f), oz = fmaxf(qz, 0.f);
        float o2 = ox * ox + oy * oy + oz * oz;
        float mx = fmaxf(qx, fmaxf(qy, qz));
        if (o2 > 0.f) {
            float inv = rsqrtf(o2);
            gx = copysignf(ox * inv, px); gy = copysignf(oy * inv, py); gz = copysignf(oz * inv, pz);
        } else {
            gx = gy = gz = 0.f;
            int imax = (qy > qx) ? 1 : 0;
            float qm = imax ? qy : qx;
            if (qz > qm) imax = 2;
            if (imax == 0) gx = px < 0.f ? -1.f : 1.f;
            else if (imax == 1) gy = py < 0.f ? -1.f : 1.f;
            else gz = pz < 0.f ? -1.f : 1.f;
        }
        return sqrtf(o2) + fminf(mx, 0.f);
    }
    if (P.mean_kind == 2) {
        gx = P.pn[0]; gy = P.pn[1]; gz = P.pn[2];
        return P.pn[0] * x + P.pn[1] * y + P.pn[2] * z - P.poff;
    }
    float px = x - P.mc[0], py = y - P.mc[1], pz = z - P.mc[2];
    float r = sqrtf(px * px + py * py + pz * pz);
    float inv = 1.0f / r;
    gx = px * inv; gy = py * inv; gz = pz * inv;
    return r - P.mr;
}

// ProceduralMean (mean_kind 4) at a double point, in double; g may be null. Out of line:
// the shapes are long and every flight kernel would otherwise inline them at each call.
__device__ __noinline__ double proc_mean_d(const sdf::MeanParams *M, double x, double y, double z, double *g) {
    return sdf::mean(*M, x, y, z, g);
}
__device__ __forceinline__ double proc_mean_value(const FieldParams &P, float x, float y, float z, float &gx,
                                                  float &gy, float &gz) {
    double g[3];
    const double v = proc_mean_d(P.sdf, (double)x, (double)y, (double)z, g);
    gx = (float)g[0]; gy = (float)g[1]; gz = (float)g[2];
    return v;
}

// Mean kinds as a compile-time choice: any (runtime switch), analytic only (sphere, box,
// plane), or the evaluated means (tabulated grid or procedural SDF, switched at run time).
enum { MEAN_ANY = 0, MEAN_ANALYTIC = 1, MEAN_EXT = 2 };

// The mean in double (the evaluated means as the reference returns them; the analytic means'
// float value widened) and its gradient in float.
template <int MEAN = MEAN_ANY>
__device__ __forceinline__ double mean_value_d(const FieldParams &P, float x, float y, float z, float &gx,
                                               float &gy, float &gz) {
    if (MEAN == MEAN_ANALYTIC) return (double)analytic_mean_value(P, x, y, z, gx, gy, gz);
    if (P.mean_kind == 3) return grid_mean_value(P, x, y, z, gx, gy, gz);
    if (P.mean_kind == 4) return proc_mean_value(P, x, y, z, gx, gy, gz);
    return (double)analytic_mean_value(P, x, y, z, gx, gy, gz);
}
template <int MEAN = MEAN_ANY>
__device__ __forceinline__ float mean_value(const FieldParams &P, float x, float y, float z, float &gx, float &gy,
                                            float &gz) {
    if (MEAN == MEAN_ANALYTIC) return analytic_mean_value(P, x, y, z, gx, gy, gz);
    return (float)mean_value_d<MEAN>(P, x, y, z, gx, gy, gz);
}

}  // namespace gpis
