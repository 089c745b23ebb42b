/*
 * gpis_oracle.c — CPU restatement of the sparse-convolution GPIS ray flight.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg.
 *
 * Build: plain C11, -O2 -ffp-contract=off, no fast-math, so that every float
 * and double operation rounds as written.  Types follow the reference:
 * a float in the reference is a float here, a double a double.  Where the
 * reference calls an unqualified libm function on a float, the float overload
 * is used (logf, powf, floorf); exp on float kernel arguments is evaluated in
 * double and rounded, which is within 1 ulp of either overload.
 *
 * Citations are file:line in the reference tree.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_FP64_TRUTH
/* The FP64 truth build (oracle/Makefile target fp64): the same algorithm with every float
 * of the field arithmetic a double and every float libm call its double version, so that
 * the only remaining rounding is double's.  What the reference defines as binary32 data
 * keeps its value: the generator's draws (u01 below), float constants of the reference
 * and the request/reply structs of the header (orc_f32).  Pointer arguments that the
 * header declares as float are double in this build. */
#define float double
#define sqrtf sqrt
#define floorf floor
#define fmaxf fmax
#define fminf fmin
#define logf log
#define powf pow
#define sinf sin
#define cosf cos
#define copysignf copysign
#define fabsf fabs
#endif

#include "gpis_oracle.h"

#define PI_F 3.1415926536f /* src/core/math/Angle.hpp:8 (float) */

typedef struct { float x, y, z; } v3f;
typedef struct { double x, y, z; } v3d;
typedef struct { float x, y, z, w; } v4f;
typedef struct { float x, y; } v2f;
typedef struct { float m[3][3]; } m3f; /* m[row][col] */

static inline v3f V3(float x, float y, float z) { v3f r = {x, y, z}; return r; }
static inline v3f vadd(v3f a, v3f b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3f vsub(v3f a, v3f b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3f vmul(v3f a, v3f b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3f vscale(float s, v3f a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline v3f vdivs(v3f a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
/* Vec::dot: sum = a0*b0; sum += a1*b1; sum += a2*b2 (Vec.hpp:200-206) */
static inline float vdot(v3f a, v3f b) { float s = a.x * b.x; s += a.y * b.y; s += a.z * b.z; return s; }
/* Vec::lengthSq starts from 0 (Vec.hpp:134-140) */
static inline float vlen2(v3f a) { float s = 0.f; s += a.x * a.x; s += a.y * a.y; s += a.z * a.z; return s; }
static inline float vlen(v3f a) { return sqrtf(vlen2(a)); }
/* Vec::normalized multiplies by 1/length (Vec.hpp:191-198) */
static inline v3f vnorm(v3f a) { float inv = 1.0f / vlen(a); return V3(a.x * inv, a.y * inv, a.z * inv); }
static inline v3f vcross(v3f a, v3f b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v4f V4(float x, float y, float z, float w) { v4f r = {x, y, z, w}; return r; }
static inline v4f v4add(v4f a, v4f b) { return V4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4f v4scale(float s, v4f a) { return V4(s * a.x, s * a.y, s * a.z, s * a.w); }
static inline v4f v4divs(v4f a, float s) { return V4(a.x / s, a.y / s, a.z / s, a.w / s); }
static inline v3d D3(double x, double y, double z) { v3d r = {x, y, z}; return r; }
static inline v3d d_of(v3f a) { return D3(a.x, a.y, a.z); }
static inline v3f f_of(v3d a) { return V3((float)a.x, (float)a.y, (float)a.z); }
static inline double ddot(v3d a, v3d b) { double s = a.x * b.x; s += a.y * b.y; s += a.z * b.z; return s; }
static inline double dlen2(v3d a) { double s = 0.0; s += a.x * a.x; s += a.y * a.y; s += a.z * a.z; return s; }
static inline v3d dnorm(v3d a) { double inv = 1.0 / sqrt(dlen2(a)); return D3(a.x * inv, a.y * inv, a.z * inv); }
static inline float sqrf(float a) { return a * a; }

/* ===================================================================== */
/* Leaves                                                                */
/* ===================================================================== */

#define P2 2246822519U
#define P3 3266489917U
#define P4 668265263U
#define P5 374761393U
static inline uint32_t rotl17(uint32_t h) { return (h << 17) | (h >> 15); }
static inline uint32_t xxfinal(uint32_t h) {
    h = P2 * (h ^ (h >> 15));
    h = P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
/* MathUtil.hpp:179-187 */
uint32_t orc_xxhash32_1(uint32_t p) { uint32_t h = p + P5; h = P4 * rotl17(h); return xxfinal(h); }
/* MathUtil.hpp:190-198: key (x, y) = (p.x, p.y) */
uint32_t orc_xxhash32_2(uint32_t x, uint32_t y) {
    uint32_t h = y + P5 + x * P3;
    h = P4 * rotl17(h);
    return xxfinal(h);
}
/* MathUtil.hpp:212-224 */
uint32_t orc_xxhash32_4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    uint32_t h = w + P5 + x * P3;
    h = P4 * rotl17(h);
    h += y * P3; h = P4 * rotl17(h);
    h += z * P3; h = P4 * rotl17(h);
    return xxfinal(h);
}

/* UniformSampler (UniformSampler.hpp:17-75), sequence 0 => increment 1. */
typedef struct { uint64_t s; } pcg_t;
static inline uint32_t pcg_next(pcg_t *g) {
    uint64_t old = g->s;
    g->s = old * 6364136223846793005ULL + 1ULL;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((uint32_t)(-(int32_t)rot) & 31u));
}
/* BitManip::normalizedUint (BitManip.hpp:47-50) */
static inline float u01(uint32_t i) {
    union { uint32_t u; orc_f32 f; } c; c.u = (i >> 9u) | 0x3F800000u; return c.f - 1.0f;
}
static inline float pcg_1d(pcg_t *g) { return u01(pcg_next(g)); }
/* set_state: state = seed, then next2D() discards two draws (UniformSampler.hpp:71-75) */
static inline void pcg_set(pcg_t *g, uint64_t st) { g->s = st; (void)pcg_next(g); (void)pcg_next(g); }
/* Default constructor seeds 0xBA5EBA11 (UniformSampler.hpp:17-27). */
static inline void pcg_default(pcg_t *g) { pcg_set(g, 0xBA5EBA11ULL); }

/* next3D() = Vec3f(next1D(), next1D(), next1D()) (UniformSampler.hpp:62-64).
 * GCC evaluates these constructor arguments right to left: the first draw is z
 * and the third is x (pinned against the reference headers built by g++). */
static inline v3f pcg_3d(pcg_t *g) { float z = pcg_1d(g); float y = pcg_1d(g); float x = pcg_1d(g); return V3(x, y, z); }

void orc_cell3d_impulses(uint32_t i, uint32_t j, uint32_t k, uint32_t seed, uint32_t n, float *out) {
    pcg_t g;
    /* cell3D: set_state(xxhash32(Vec4u(ijk.z, ijk.y, ijk.x, seed)) + 1u) (SparseConvolutionNoise.cpp:433) */
    pcg_set(&g, (uint64_t)(uint32_t)(orc_xxhash32_4(k, j, i, seed) + 1u));
    for (uint32_t q = 0; q < n; ++q) {
        v3f p = pcg_3d(&g);
        float w = pcg_1d(&g) < 0.5f ? -1.f : 1.f; /* Bernoulli (MathUtil.hpp:226-228) */
        out[4 * q + 0] = p.x; out[4 * q + 1] = p.y; out[4 * q + 2] = p.z; out[4 * q + 3] = w;
    }
}

void orc_cell1d_impulses(uint32_t i, uint32_t seed, uint32_t n, float *out) {
    pcg_t g;
    pcg_set(&g, (uint64_t)(uint32_t)(orc_xxhash32_2(i, seed) + 1u)); /* :466 */
    for (uint32_t q = 0; q < n; ++q) {
        float t = pcg_1d(&g);
        float w = pcg_1d(&g) < 0.5f ? -1.f : 1.f;
        out[2 * q] = t; out[2 * q + 1] = w;
    }
}

/* rand_normal_2 (Gaussian.cpp:22-34): 2*PI is float, the rest double. */
static void box_muller(pcg_t *g, double *z1, double *z2) {
    double a = pcg_1d(g);
    double b = pcg_1d(g);
    double r = sqrt(-2 * log(1. - a));
    double two_pi = (double)(2 * PI_F);
    double x = cos(two_pi * b);
    double y = sin(two_pi * b);
    *z1 = r * x; *z2 = r * y;
}
void orc_standard_normals(uint64_t state, int npairs, double *out) {
    pcg_t g; pcg_set(&g, state);
    for (int m = 0; m < npairs; ++m) box_muller(&g, &out[2 * m], &out[2 * m + 1]);
}

/* TangentFrame(n) (TangentFrame.hpp:74-85) */
typedef struct { v3f normal, tangent, bitangent; } frame_t;
static frame_t make_frame(v3f n) {
    frame_t F;
    F.normal = vnorm(n);
    float sign = copysignf(1.0f, F.normal.z);
    const float a = -1.0f / (sign + F.normal.z);
    const float b = F.normal.x * F.normal.y * a;
    F.tangent = V3(1.0f + sign * F.normal.x * F.normal.x * a, sign * b, -sign * F.normal.x);
    F.bitangent = V3(b, sign + F.normal.y * F.normal.y * a, -F.normal.y);
    F.tangent = vnorm(F.tangent);
    F.bitangent = vnorm(F.bitangent);
    return F;
}
static inline v3f to_local(const frame_t *F, v3f p) { return V3(vdot(F->tangent, p), vdot(F->bitangent, p), vdot(F->normal, p)); }
static inline v3f to_global(const frame_t *F, v3f p) {
    return vadd(vadd(vscale(p.x, F->tangent), vscale(p.y, F->bitangent)), vscale(p.z, F->normal));
}
void orc_tangent_frame(const float n[3], float out9[9]) {
    frame_t F = make_frame(V3(n[0], n[1], n[2]));
    out9[0] = F.tangent.x; out9[1] = F.tangent.y; out9[2] = F.tangent.z;
    out9[3] = F.bitangent.x; out9[4] = F.bitangent.y; out9[5] = F.bitangent.z;
    out9[6] = F.normal.x; out9[7] = F.normal.y; out9[8] = F.normal.z;
}

/* 3x3 float inverse by cofactors and 1/det (Eigen compute_inverse size 3). */
static m3f m3_inverse(m3f A) {
    float (*m)[3] = A.m;
    float c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    float c10 = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    float c20 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    float det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    float inv = 1.0f / det;
    m3f R;
    R.m[0][0] = c00 * inv; R.m[0][1] = c10 * inv; R.m[0][2] = c20 * inv;
    R.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
    R.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    R.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    R.m[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
    R.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    R.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    return R;
}
static v3f m3_mul(const m3f *A, v3f v) {
    return V3(A->m[0][0] * v.x + A->m[0][1] * v.y + A->m[0][2] * v.z,
              A->m[1][0] * v.x + A->m[1][1] * v.y + A->m[1][2] * v.z,
              A->m[2][0] * v.x + A->m[2][1] * v.y + A->m[2][2] * v.z);
}

/* ===================================================================== */
/* Field: covariance, mean, realization                                  */
/* ===================================================================== */

struct orc_field {
    orc_medium_desc d;
    float l_conv;          /* _l_conv = _l*sqrt(2.f)/2 (GPFunctions.cpp:661) */
    v3f l_aniso, l_aniso_inv;
    int multires;          /* SparseConvolutionNoise.cpp:47 */
    /* useAnisoMtx (GPFunctions.cpp:675-681): _local_to_world = l_conv * anisoMtx, its float
     * inverse, _cov_mtx_inv = W^T W and its determinant */
    int full;
    float l2w[3][3], w2l[3][3], cinv[3][3], cinv_det;
    int activate;          /* :38 */
    int scheme1d;          /* :40-43 */
    float base;            /* :49 */
};

typedef struct {
    float value_scale; v3f grad_scale; v3f origin;                 /* coeff_3D */
    float v1; v3f g1; v3f o1; v3f target_grad; int has_grad;     /* coeff_1D */
} cond_t;

typedef struct {
    const orc_field *f;
    uint32_t seed;         /* computeSeed(info) */
    uint32_t normal_seed;  /* xxh(pss) + xxh(sceneSeed) + 1 */
    cond_t c;
    float perturb;         /* test hook: +-perturb added to each marched value (0 = off) */
    uint32_t pstate;
} real_t;

static int is_se(const orc_field *f) { return f->d.cov_kind == ORC_COV_SE; }

/* ---- ProceduralNoise fields over SdfFunctions' simplex fbm ----------------------
 * random3, simplex3d, fbm: SdfFunctions.cpp:84-183 (float arithmetic as written; random3's
 * sin takes the float dot in double, then 4096 * sin rounds to float -- the reference
 * build resolves the unqualified sin to the double one). */
static const float ORC_F3 = 0.3333333f, ORC_G3 = 0.1666667f;
static v3f orc_random3_v(v3f c) {
    float dd = vdot(c, V3(17.0f, 59.4f, 15.0f));
    float j = (float)(4096.0 * sin((double)dd));
    double t = 512.0 * (double)j;
    float rz = (float)(t - floor(t));
    j = (float)((double)j * 0.125);
    t = 512.0 * (double)j;
    float rx = (float)(t - floor(t));
    j = (float)((double)j * 0.125);
    t = 512.0 * (double)j;
    float ry = (float)(t - floor(t));
    return V3(rx - 0.5f, ry - 0.5f, rz - 0.5f);
}
static float orc_simplex3d_v(v3f p) {
    float dp = vdot(p, V3(ORC_F3, ORC_F3, ORC_F3));
    v3f s = V3(floorf(p.x + dp), floorf(p.y + dp), floorf(p.z + dp));
    float ds = vdot(s, V3(ORC_G3, ORC_G3, ORC_G3));
    v3f x = V3((p.x - s.x) + ds, (p.y - s.y) + ds, (p.z - s.z) + ds);
    /* e = step(0, x - x.yzx); i1 = e * (1 - e.zxy); i2 = 1 - e.zxy * (1 - e) */
    v3f e = V3((x.x - x.y) < 0.0f ? 0.0f : 1.0f, (x.y - x.z) < 0.0f ? 0.0f : 1.0f, (x.z - x.x) < 0.0f ? 0.0f : 1.0f);
    v3f ezxy = V3(e.z, e.x, e.y);
    v3f i1 = V3(e.x * (1.0f - ezxy.x), e.y * (1.0f - ezxy.y), e.z * (1.0f - ezxy.z));
    v3f i2 = V3(1.0f - ezxy.x * (1.0f - e.x), 1.0f - ezxy.y * (1.0f - e.y), 1.0f - ezxy.z * (1.0f - e.z));
    const float g2 = 2.0f * ORC_G3, g3 = 3.0f * ORC_G3;
    v3f x1 = V3((x.x - i1.x) + ORC_G3, (x.y - i1.y) + ORC_G3, (x.z - i1.z) + ORC_G3);
    v3f x2 = V3((x.x - i2.x) + g2, (x.y - i2.y) + g2, (x.z - i2.z) + g2);
    v3f x3 = V3((x.x - 1.0f) + g3, (x.y - 1.0f) + g3, (x.z - 1.0f) + g3);
    float w[4] = {vdot(x, x), vdot(x1, x1), vdot(x2, x2), vdot(x3, x3)};
    for (int i = 0; i < 4; ++i) { float a = 0.6f - w[i]; w[i] = a < 0.0f ? 0.0f : a; }
    float d[4];
    d[0] = vdot(orc_random3_v(s), x);
    d[1] = vdot(orc_random3_v(V3(s.x + i1.x, s.y + i1.y, s.z + i1.z)), x1);
    d[2] = vdot(orc_random3_v(V3(s.x + i2.x, s.y + i2.y, s.z + i2.z)), x2);
    d[3] = vdot(orc_random3_v(V3(s.x + 1.0f, s.y + 1.0f, s.z + 1.0f)), x3);
    for (int i = 0; i < 4; ++i) { w[i] *= w[i]; }
    for (int i = 0; i < 4; ++i) { w[i] *= w[i]; }
    for (int i = 0; i < 4; ++i) { d[i] *= w[i]; }
    float r = d[0] * 52.0f;
    r += d[1] * 52.0f;
    r += d[2] * 52.0f;
    r += d[3] * 52.0f;
    return r;
}
static double orc_fbm_v(v3d uv, int octaves) {
    const float gain = (float)0.65, lacunarity = (float)2.1042;
    float frequency = 0.5f, amplitude = gain;
    uv = D3(uv.x * 5.0, uv.y * 5.0, uv.z * 5.0);
    v3f u = f_of(uv);
    float total = orc_simplex3d_v(u);
    for (int i = 0; i < octaves; ++i) {
        total += orc_simplex3d_v(V3(u.x * frequency, u.y * frequency, u.z * frequency)) * amplitude;
        frequency *= lacunarity;
        amplitude *= gain;
    }
    total = (float)(((double)total + 2.0) / 4.0);
    return (double)total;
}
static double dclamp(double v, double lo, double hi) { double a = v < lo ? lo : v; return a > hi ? hi : a; }
static double dlerp(double a, double b, double t) { return a * (1.0 - t) + b * t; }
/* sqrt(exp(lerp(log(min^2), log(max^2), clamp(c * scale + offset, 0, 1)))) */
static double noise_ramp(double mn, double mx, double c, double scale, double offset) {
    return sqrt(exp(dlerp(log(mn * mn), log(mx * mx), dclamp(c * scale + offset, 0.0, 1.0))));
}
/* ProceduralNoise::operator() (vec = 0) and the largest component of
 * ProceduralNoiseVec::operator() (vec = 1), GPFunctions.cpp:53-119; q = min, max, start,
 * end, min2, max2, start2, end2; _const = 1; scale = 1 / (end - start), offset = -start * scale. */
static double noise_field(int type, const double *q, int vec, v3d p) {
    const double cst = 1.0;
    const double scale = 1.0 / (q[3] - q[2]), offset = -q[2] * scale;
    const double scale2 = 1.0 / (q[7] - q[6]), offset2 = -q[6] * scale2;
    const double mn = q[0] + cst, mx = q[1] + cst;
    double v;
    switch (type) {
    case ORC_NOISE_BOTTOM_TOP: v = noise_ramp(mn, mx, p.y, scale, offset) - cst; break;
    case ORC_NOISE_LEFT_RIGHT: v = noise_ramp(mn, mx, p.x, scale, offset) - cst; break;
    case ORC_NOISE_FRONT_BACK: v = noise_ramp(mn, mx, p.z, scale, offset) - cst; break;
    case ORC_NOISE_BOTTOM_TOP_LEFT_RIGHT: {
        const double mn2 = q[4] + cst, mx2 = q[5] + cst;
        float bt = (float)noise_ramp(mn, mx, p.y, scale, offset);
        float lr = (float)noise_ramp(mn2, mx2, p.x, scale2, offset2);
        v = (double)(bt * lr) - cst * cst;
        break;
    }
    case ORC_NOISE_SANDSTONE: {
        const int oct = vec ? 10 : 2;
        v3d ps = D3(p.x * 0.3, p.y * 0.3, p.z * 0.3);
        double a = orc_fbm_v(ps, oct);
        double b = orc_fbm_v(D3(ps.x + a, ps.y + a, ps.z + a), oct);
        double f = orc_fbm_v(D3(ps.x + b, ps.y + b, ps.z + b), oct);
        if (!vec) {
            double cx = sqrt(f * 1.9 * 1.2) - 0.35;
            v = dlerp(q[0], q[1], dclamp(cx, 0., 1.));
        } else {
            /* clamp(col * 0.2, 0, 1) per component; the largest as floats */
            float cx = (float)dclamp((sqrt(f * 1.9 * 1.2) - 0.35) * 0.2, 0., 1.);
            float cy = (float)dclamp((sqrt(f * 0.7 * 1.2) - 0.35) * 0.2, 0., 1.);
            float cz = (float)dclamp((sqrt(f * 0.25 * 1.2) - 0.35) * 0.2, 0., 1.);
            return (double)fmaxf(fmaxf(cx, cy), cz);
        }
        break;
    }
    case ORC_NOISE_RUST: default: {
        const int oct = vec ? 10 : 2;
        v3d ps = D3(p.x * 2.0, p.y * 2.0, p.z * 2.0);
        double inner = orc_fbm_v(D3(ps.x * .1, ps.y * .1, ps.z * .1), oct) * 0.4;
        double a = orc_fbm_v(D3(ps.x + inner, ps.y + inner, ps.z + inner), oct);
        double c = orc_fbm_v(D3(ps.x * 25., ps.y * 25., ps.z * 25.), oct) * 0.1;
        double x = vec ? a + c : a - c;
        x = dclamp((x - 0.4) / (0.6 - 0.4), 0.0, 1.0);
        double f = x * x * (3.0 - 2.0 * x);
        if (!vec) {
            v = dlerp(q[0], q[1], dclamp(f, 0., 1.));
        } else {
            float cx = (float)(0.278 * (1.0 - f) + 1.0 * f);
            float cy = (float)(0.212 * (1.0 - f) + 1.0 * f);
            float cz = (float)(0.141 * (1.0 - f) + 1.0 * f);
            return (double)fmaxf(fmaxf(cx, cy), cz);
        }
        break;
    }
    }
    return vec ? (double)(float)v : v;
}
void orc_random3(const float *p, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        v3f r = orc_random3_v(V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}
void orc_simplex3d(const float *p, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_simplex3d_v(V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
}
void orc_fbm(const double *p, int64_t n, int octaves, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_fbm_v(D3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), octaves);
}
void orc_noise_field(int32_t type, const double params8[8], int vec, const double *p, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = noise_field(type, params8, vec, D3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
}

/* ---- procedural nonstationary fields (D3: defined in code) ---------- */
static double ns_ls_value(const orc_field *f, v3d p) {
    const orc_f32 *q = f->d.ns_ls_params;
    if (f->d.ns_ls_field == ORC_LS_NOISE) return noise_field(f->d.ns_ls_noise, f->d.ns_ls_noise_params, 1, p);
    switch (f->d.ns_ls_field) {
    case ORC_LS_SINCOS: {
        double fr = q[2];
        return (double)q[0] + (double)q[1] * (0.5 + 0.5 * sin(fr * p.x) * cos(fr * p.z));
    }
    default:
        return q[0];
    }
}

/* ---- CovarianceFunction virtuals (GPFunctions.cpp:610-875, 1225-1345, 1600-1760) */
static float cov_amplitude(const orc_field *f, v3f p) {
    if (is_se(f)) return f->d.sigma;                       /* :721 */
    if (f->d.ns_var_field == ORC_VAR_NOISE)                 /* getVariance(p) * sigma, :1245-1247 */
        return (float)(noise_field(f->d.ns_var_noise, f->d.ns_var_noise_params, 0, d_of(p)) * (double)f->d.sigma);
    return (float)((double)f->d.ns_var * f->d.sigma);       /* :1241-1243 */
}
static float cov_lateral(const orc_field *f, v3f p) {
    if (is_se(f)) return 1.0f;                               /* :610-612 */
    float l = (float)ns_ls_value(f, d_of(p));                /* getKernelScale :1734-1740 */
    return fmaxf(fmaxf(l, l), l);
}
static float cov_ls_max(const orc_field *f) { return is_se(f) ? 1.f : f->d.ns_ls_max; }
static float cov_splat_scale(const orc_field *f, v3f p) {  /* nonStationarySplattingKernelScale */
    if (is_se(f) || f->multires) return 1.f;
    return cov_lateral(f, p) / cov_ls_max(f);
}
static float cov_world_sampling_scale(const orc_field *f) { return cov_ls_max(f); }
static float se_kernel_radius(const orc_field *f, int identity, float local_scale) { /* :706-719 */
    float s = f->d.kernel_scale;
    if (identity) return s;
    float mf;
    if (!f->full) {
        mf = fmaxf(fmaxf(f->l_aniso.x, f->l_aniso.y), f->l_aniso.z);
    } else {
        /* end_pt = col(0) + col(1) + col(2), its largest component */
        float e[3];
        for (int i = 0; i < 3; ++i) e[i] = (f->l2w[i][0] + f->l2w[i][1]) + f->l2w[i][2];
        mf = fmaxf(fmaxf(e[0], e[1]), e[2]);
    }
    return s * local_scale * mf;
}
static float cov_kernel_radius(const orc_field *f, int identity, float local_scale) {
    if (!is_se(f)) local_scale *= f->multires ? 1.0f : cov_ls_max(f);   /* :1251-1255 */
    return se_kernel_radius(f, identity, local_scale);
}
static float se_var3d(const orc_field *f, float N, float R, int identity, float gs, float ls) { /* :751-764 */
    double dens = N / (R * R * R);
    double det = 1.0;
    if (!identity) {
        if (!f->full) det = f->l_aniso.x * f->l_aniso.y * f->l_aniso.z;
        else det = 1.0 / sqrt(f->cinv_det);
        det *= pow(gs, 3);
    }
    det *= pow(ls, 3);
    double integral = pow(M_PI, 1.5) * det;
    return (float)(dens * integral);
}
static float cov_var3d(const orc_field *f, v3f p, float N, float R, int identity, float gs) {
    if (is_se(f)) return se_var3d(f, N, R, identity, gs, 1.0f);
    float ls = f->multires ? 1.0f : cov_splat_scale(f, p);
    return se_var3d(f, N, R, identity, gs, ls);
}
static float se_var1d(float N, float R, float ls) { /* :766-770 */
    double dens = N / R;
    double integral = sqrt(M_PI) * ls;
    return (float)(dens * integral);
}
static float cov_var1d(const orc_field *f, v3f p, float N, float R) {
    if (is_se(f)) return se_var1d(N, R, 1.0f);
    return se_var1d(N, R, cov_splat_scale(f, p));
}
/* transforms (:634-644; nonstationary :1267-1285) */
static float ns_scale(const orc_field *f, float gs) { return (is_se(f) || f->multires) ? gs : cov_ls_max(f); }
/* M v and M^T v for a row-major 3x3 float matrix, each component left to right */
static v3f mat_mul(const float M[3][3], v3f v) {
    return V3((M[0][0] * v.x + M[0][1] * v.y) + M[0][2] * v.z, (M[1][0] * v.x + M[1][1] * v.y) + M[1][2] * v.z,
              (M[2][0] * v.x + M[2][1] * v.y) + M[2][2] * v.z);
}
static v3f mat_tmul(const float M[3][3], v3f v) {
    return V3((M[0][0] * v.x + M[1][0] * v.y) + M[2][0] * v.z, (M[0][1] * v.x + M[1][1] * v.y) + M[2][1] * v.z,
              (M[0][2] * v.x + M[1][2] * v.y) + M[2][2] * v.z);
}
/* transformPosDirWorldtoLocal: _world_to_local * v / s; transformGradLocaltoWorld:
 * _world_to_local^T * g / s; transformGradWorldtoLocal: _local_to_world^T * g * s
 * (GPFunctions.cpp:634-644); the diagonal matrices of "aniso" or the full "anisoMtx" */
static v3f w2l(const orc_field *f, v3f v, float gs) {
    float s = ns_scale(f, gs);
    if (f->full) return vdivs(mat_mul(f->w2l, v), s);
    return vdivs(vmul(f->l_aniso_inv, v), s);
}
static v3f grad_l2w(const orc_field *f, v3f g, float gs) {
    float s = ns_scale(f, gs);
    if (f->full) return vdivs(mat_tmul(f->w2l, g), s);
    return vdivs(vmul(f->l_aniso_inv, g), s);
}
static v3f grad_w2l(const orc_field *f, v3f g, float gs) {
    float s = ns_scale(f, gs);
    v3f r = f->full ? mat_tmul(f->l2w, g) : vmul(f->l_aniso, g);
    return V3(r.x * s, r.y * s, r.z * s);
}

/* getInvCovMtx diagonal (GPFunctions.cpp:784-812), no anisotropy field. */
static v3f inv_cov_diag(const orc_field *f, int is_cov, int is_iso, float gs, float ls) {
    v3f m;
    if (!is_iso) {
        m = vmul(f->l_aniso_inv, f->l_aniso_inv);
        float g2 = sqrf(gs);
        m = V3(m.x / g2, m.y / g2, m.z / g2);
    } else {
        m = V3(1.f, 1.f, 1.f);
    }
    if (is_cov) m = vscale(0.5f, m);
    float l2 = sqrf(ls);
    m = V3(m.x / l2, m.y / l2, m.z / l2);
    return vscale(0.5f, m);
}
/* getInvCovMtx with the full matrix (useAnisoMtx, not isotropic): _cov_mtx_inv / gs^2,
 * times 0.5 for a covariance, / ls^2, times 0.5 (GPFunctions.cpp:784-812) */
static m3f inv_cov_full(const orc_field *f, int is_cov, float gs, float ls) {
    m3f m;
    float g2 = sqrf(gs), l2 = sqrf(ls);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float v = f->cinv[i][j] / g2;
            if (is_cov) v *= 0.5f;
            v /= l2;
            m.m[i][j] = v * 0.5f;
        }
    return m;
}
/* ab^T M ab (dist2_ab, GPFunctions.hpp:77-80) and -2 ab . M.col(i) */
static float quad_form(const m3f *m, v3f ab, v3f *dots) {
    float r0 = ab.x * m->m[0][0] + ab.y * m->m[1][0] + ab.z * m->m[2][0];
    float r1 = ab.x * m->m[0][1] + ab.y * m->m[1][1] + ab.z * m->m[2][1];
    float r2 = ab.x * m->m[0][2] + ab.y * m->m[1][2] + ab.z * m->m[2][2];
    if (dots) *dots = V3(-2.f * r0, -2.f * r1, -2.f * r2);
    return (r0 * ab.x + r1 * ab.y) + r2 * ab.z;
}
/* CovarianceFunction::splattingKernel3D (GPFunctions.cpp:555-565, 814-827) */
static v4f splat3d(const orc_field *f, v3f pa, v3f pb, int is_cov, int is_iso, float gs, v3f p_world) {
    float ls = cov_splat_scale(f, p_world);
    if (f->full && !is_iso) {
        m3f M = inv_cov_full(f, is_cov, gs, ls);
        v3f dots;
        float absq = quad_form(&M, vsub(pa, pb), &dots);
        float v = (float)exp(-(double)absq);
        return V4(v, dots.x * v, dots.y * v, dots.z * v);
    }
    v3f m = inv_cov_diag(f, is_cov, is_iso, gs, ls);
    v3f ab = vsub(pa, pb);
    float absq = ab.x * m.x * ab.x + ab.y * m.y * ab.y + ab.z * m.z * ab.z;
    float v = (float)exp(-(double)absq);
    return V4(v, -2.f * (ab.x * m.x) * v, -2.f * (ab.y * m.y) * v, -2.f * (ab.z * m.z) * v);
}
/* CovarianceFunction::splattingKernel3DGrad (:567-580, 829-843) */
static v4f splat3d_grad(const orc_field *f, v3f pa, v3f pb, v3f coeff, int is_cov, int is_iso, float gs, v3f p_world) {
    float ls = cov_splat_scale(f, p_world);
    if (f->full && !is_iso) {
        m3f M = inv_cov_full(f, is_cov, gs, ls);
        v3f d;
        float absq = quad_form(&M, vsub(pa, pb), &d);
        float v = (float)exp(-(double)absq);
        float dd[3] = {d.x, d.y, d.z};
        v4f rows[3];
        for (int i = 0; i < 3; ++i) {
            float h[3];
            for (int j = 0; j < 3; ++j) h[j] = (dd[i] * dd[j] - 2.f * M.m[i][j]) * v;
            rows[i] = V4(dd[i] * v, h[0], h[1], h[2]);
        }
        return v4add(v4add(v4scale(coeff.x, rows[0]), v4scale(coeff.y, rows[1])), v4scale(coeff.z, rows[2]));
    }
    v3f m = inv_cov_diag(f, is_cov, is_iso, gs, ls);
    v3f ab = vsub(pa, pb);
    float absq = ab.x * m.x * ab.x + ab.y * m.y * ab.y + ab.z * m.z * ab.z;
    float v = (float)exp(-(double)absq);
    float dx = -2.f * (ab.x * m.x), dy = -2.f * (ab.y * m.y), dz = -2.f * (ab.z * m.z);
    float g1x = dx * v, g1y = dy * v, g1z = dz * v;
    float h00 = (dx * dx - 2.f * m.x) * v, h11 = (dy * dy - 2.f * m.y) * v, h22 = (dz * dz - 2.f * m.z) * v;
    float h01 = (dx * dy) * v, h02 = (dx * dz) * v, h12 = (dy * dz) * v;
    v4f vx = V4(g1x, h00, h01, h02), vy = V4(g1y, h01, h11, h12), vz = V4(g1z, h02, h12, h22);
    return v4add(v4add(v4scale(coeff.x, vx), v4scale(coeff.y, vy)), v4scale(coeff.z, vz));
}
/* sparseConvNoiseOneOverSecondDerivative (:723-749), no anisotropy field */
static m3f one_over_second(const orc_field *f, int is_iso) {
    m3f M; memset(&M, 0, sizeof M);
    if (f->full && !is_iso) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M.m[i][j] = -2.f * (f->cinv[i][j] * 0.25f);
        return m3_inverse(M);
    }
    v3f d = is_iso ? V3(1.f, 1.f, 1.f) : vmul(f->l_aniso_inv, f->l_aniso_inv);
    M.m[0][0] = d.x * 0.25f * -2.f; M.m[1][1] = d.y * 0.25f * -2.f; M.m[2][2] = d.z * 0.25f * -2.f;
    return m3_inverse(M);
}
/* 1D kernels (GPFunctions.cpp:845-875) */
static float se_splat1d(float ab, float ls) { float den = (float)(2.0 * sqrf(ls)); return (float)exp(-(double)(ab * ab / den)); }
static float se_splat1d_g(float ab, float ls) {
    float fv = se_splat1d(ab, ls);
    float den = (float)(2.0 * sqrf(ls));
    return (float)(-2.0 * ab / den * fv);
}
static float sqrt2ls(float ls) { return (float)(ls * sqrt(2.0)); }
static float se_cov1d(float ab, float ls) { return se_splat1d(ab, sqrt2ls(ls)); }
static float se_cov1d_g(float ab, float ls) { return se_splat1d_g(ab, sqrt2ls(ls)); }
static float se_cov1d_gg(float ab, float ls) {
    float fv = se_cov1d(ab, ls);
    float den = (float)(4.0 * sqrf(ls));
    float g1 = (float)(-2.0 * ab / den);
    return (float)((g1 * g1 - 2.0 / den) * fv);
}
static float se_cov1d_gg3(float ab, float ls) {
    float fv = se_cov1d(ab, ls);
    float den = (float)(4.0 * sqrf(ls));
    return (float)(-2.0 / den * fv);
}
/* getNonstationaryCovSplatCov1D (:1720-1732), no anisotropy field */
static float ns_cov_splat_cov1d(const orc_field *f, v3f pq, v3f pc) {
    if (is_se(f)) return 1.0f;
    float s = (float)(0.5 * (sqrf(cov_splat_scale(f, pq)) + sqrf(cov_splat_scale(f, pc))));
    return sqrtf(s);
}
/* CovarianceFunction::splattingKernel1D (:582-588) */
static v2f splat1d(const orc_field *f, float pq, float pc, v3f pcw) {
    float ls = cov_splat_scale(f, pcw);
    ls *= sqrtf(1.0f);
    v2f r;
    if (is_se(f)) { r.x = se_splat1d(pq - pc, 1.0f); r.y = se_splat1d_g(pq - pc, 1.0f); }
    else { r.x = se_splat1d(pq - pc, ls); r.y = se_splat1d_g(pq - pc, ls); }
    return r;
}
static v2f cov1d(const orc_field *f, float pq, float pc, v3f pqw, v3f pcw) {
    float ls = ns_cov_splat_cov1d(f, pqw, pcw);
    v2f r = { se_cov1d(pq - pc, ls), se_cov1d_g(pq - pc, ls) };
    return r;
}
static v2f cov1d_grad(const orc_field *f, float pq, float pc, v3f pqw, v3f pcw) {
    float ls = ns_cov_splat_cov1d(f, pqw, pcw);
    v2f r = { se_cov1d_g(pq - pc, ls), se_cov1d_gg(pq - pc, ls) };
    return r;
}
static float cov1d_grad3n(const orc_field *f, float pq, float pc, v3f pqw, v3f pcw) {
    return se_cov1d_gg3(pq - pc, ns_cov_splat_cov1d(f, pqw, pcw));
}

/* ---- tabulated mean on a dense grid --------------------------------- */
/* voxel(): index relative to the grid origin, clamped to the grid (RaySegmentBatch.cu:64-71) */
static float grid_voxel(const orc_medium_desc *d, int x, int y, int z) {
    x -= d->grid_origin[0]; y -= d->grid_origin[1]; z -= d->grid_origin[2];
    x = x < 0 ? 0 : (x > d->grid_dims[0] - 1 ? d->grid_dims[0] - 1 : x);
    y = y < 0 ? 0 : (y > d->grid_dims[1] - 1 ? d->grid_dims[1] - 1 : y);
    z = z < 0 ? 0 : (z > d->grid_dims[2] - 1 ? d->grid_dims[2] - 1 : z);
    return d->grid[((size_t)z * (size_t)d->grid_dims[1] + (size_t)y) * (size_t)d->grid_dims[0] + (size_t)x];
}
static float fclampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
/* meanValue()'s density: world -> index affine map, clamp to [clampMin, clampMax] (the
 * VdbGrid::density clamp to bounds+2 / bounds-3, VdbGrid.cpp:413 and :458-459), then
 * trilinear interpolation along x, then y, then z (RaySegmentBatch.cu:73-95). */
float orc_grid_density(const orc_medium_desc *d, const float p[3]) {
    const orc_f32 *m = d->grid_w2i;
    float q[3];
    for (int r = 0; r < 3; ++r) q[r] = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
    for (int i = 0; i < 3; ++i) q[i] = fclampf(q[i], d->grid_clamp_min[i], d->grid_clamp_max[i]);
    int x = (int)floorf(q[0]), y = (int)floorf(q[1]), z = (int)floorf(q[2]);
    float fx = q[0] - (float)x, fy = q[1] - (float)y, fz = q[2] - (float)z;
    float layer[2];
    for (int dz = 0; dz < 2; ++dz) {
        float row[2];
        for (int dy = 0; dy < 2; ++dy) {
            float a = grid_voxel(d, x, y + dy, z + dz);
            float b = grid_voxel(d, x + 1, y + dy, z + dz);
            row[dy] = a + fx * (b - a);
        }
        layer[dz] = row[0] + fy * (row[1] - row[0]);
    }
    return layer[0] + fz * (layer[1] - layer[0]);
}
/* TabulatedMean::mean (GPFunctions.cpp:165-172): the point rounded to float, the density
 * widened to double, the "volume" option's -log(max(1e-4, .)), then (res + offset) * scale. */
double orc_grid_mean(const orc_medium_desc *d, const double a[3]) {
    float p[3] = {(float)a[0], (float)a[1], (float)a[2]};
    double res = (double)orc_grid_density(d, p);
    if (d->grid_volume) res = -log(res > 0.0001 ? res : 0.0001);
    return (res + (double)d->grid_offset) * (double)d->grid_scale;
}
/* TabulatedMean::dmean_da (:179-190): forward differences with eps = 0.001 in double. */
void orc_grid_mean_grad(const orc_medium_desc *d, const double a[3], double g[3]) {
    const double eps = 0.001;
    double c = orc_grid_mean(d, a);
    for (int i = 0; i < 3; ++i) {
        double b[3] = {a[0], a[1], a[2]};
        b[i] = a[i] + eps;
        g[i] = (orc_grid_mean(d, b) - c) / eps;
    }
}

/* ---- means (GPFunctions.hpp:952-996 and code-defined box/plane) ----- */
static double mean_value(const orc_field *f, v3d p) {
    const orc_medium_desc *d = &f->d;
    switch (d->mean_kind) {
    case ORC_MEAN_GRID: {
        double a[3] = {p.x, p.y, p.z};
        return orc_grid_mean(d, a);
    }
    case ORC_MEAN_PROC: {
        double a[3] = {p.x, p.y, p.z};
        return orc_sdf_mean(d, a, NULL);
    }
    case ORC_MEAN_BOX: {
        double q[3], o2 = 0.0, mx = -1e300;
        double pp[3] = {p.x - d->mean_center[0], p.y - d->mean_center[1], p.z - d->mean_center[2]};
        for (int i = 0; i < 3; ++i) {
            q[i] = fabs(pp[i]) - d->mean_half[i];
            double c = q[i] > 0 ? q[i] : 0.0;
            o2 += c * c;
            if (q[i] > mx) mx = q[i];
        }
        return sqrt(o2) + (mx < 0 ? mx : 0.0);
    }
    case ORC_MEAN_PLANE:
        return d->plane_normal[0] * p.x + d->plane_normal[1] * p.y + d->plane_normal[2] * p.z - d->plane_offset;
    default: {
        v3d c = D3(p.x - d->mean_center[0], p.y - d->mean_center[1], p.z - d->mean_center[2]);
        return sqrt(dlen2(c)) - (double)d->mean_radius;
    }
    }
}
static v3d mean_grad(const orc_field *f, v3d p) {
    const orc_medium_desc *d = &f->d;
    switch (d->mean_kind) {
    case ORC_MEAN_GRID: {
        double a[3] = {p.x, p.y, p.z}, g[3];
        orc_grid_mean_grad(d, a, g);
        return D3(g[0], g[1], g[2]);
    }
    case ORC_MEAN_PROC: {
        double a[3] = {p.x, p.y, p.z}, g[3];
        orc_sdf_mean(d, a, g);
        return D3(g[0], g[1], g[2]);
    }
    case ORC_MEAN_BOX: {
        double pp[3] = {p.x - d->mean_center[0], p.y - d->mean_center[1], p.z - d->mean_center[2]};
        double q[3], g[3] = {0, 0, 0}, o2 = 0.0;
        int imax = 0;
        for (int i = 0; i < 3; ++i) {
            q[i] = fabs(pp[i]) - d->mean_half[i];
            if (q[i] > q[imax]) imax = i;
            if (q[i] > 0) o2 += q[i] * q[i];
        }
        if (o2 > 0) {
            double inv = 1.0 / sqrt(o2);
            for (int i = 0; i < 3; ++i) g[i] = (q[i] > 0 ? q[i] : 0.0) * inv * (pp[i] < 0 ? -1.0 : 1.0);
        } else {
            g[imax] = pp[imax] < 0 ? -1.0 : 1.0;
        }
        return D3(g[0], g[1], g[2]);
    }
    case ORC_MEAN_PLANE:
        return D3(d->plane_normal[0], d->plane_normal[1], d->plane_normal[2]);
    default:
        return dnorm(D3(p.x - d->mean_center[0], p.y - d->mean_center[1], p.z - d->mean_center[2]));
    }
}

/*
 * The SDF shapes below (the knob, two-spheres, slab and groove programs) follow the shape
 * formulas of the reference's src/core/math/SdfFunctions.cpp, which carries these notices:
 *
 *   Copyright 2020 Towaki Takikawa @yongyuanxi
 *   The MIT License
 *
 *   The MIT License (MIT)
 *   Copyright (c) 2021, NVIDIA CORPORATION.
 *   Permission is hereby granted, free of charge, to any person obtaining a copy of
 *   this software and associated documentation files (the "Software"), to deal in
 *   the Software without restriction, including without limitation the rights to
 *   use, copy, modify, merge, publish, distribute, sublicense, and/or sell copies of
 *   the Software, and to permit persons to whom the Software is furnished to do so,
 *   subject to the following conditions:
 *   The above copyright notice and this permission notice shall be included in all
 *   copies or substantial portions of the Software.
 *   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
 *   IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY, FITNESS
 *   FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE AUTHORS OR
 *   COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER
 *   IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN
 *   CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
 *
 */
/* ---- procedural SDF means (GPFunctions.cpp:250-298, ProceduralMeanSdf.cpp:11-38,
 *      SdfFunctionsQueries.cpp:14-466) ---------------------------------
 * The shapes are evaluated in one of two modes, as the reference instantiates its CSG
 * templates twice: point jets (value, gradient and the constant/affine flags of
 * PointJet, :127-187) for the functions with queries, plain doubles (the float API,
 * :374-394) for knob_inner and knob_outer. */
typedef struct { double v, g[3]; int cst, aff; } sj;
typedef struct { int jet; float rp90[9], rm90[9], rm45[9]; } sdf_mode;

/* std::hypot(x, y, z) of libstdc++ 11: scale by the largest magnitude. */
static double sdf_hypot3(double x, double y, double z) {
    double m;
    x = fabs(x); y = fabs(y); z = fabs(z);
    m = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (m == 0.0) return 0.0;
    return m * sqrt((x / m) * (x / m) + (y / m) * (y / m) + (z / m) * (z / m));
}
/* guard(v) = 128 eps (1 + |v|) (SdfFunctionsQueries.cpp:15) */
static double sdf_guard(double v) { return 128.0 * 2.220446049250313e-16 * (1.0 + fabs(v)); }
static sj sj_c(double v) { sj r; r.v = v; r.g[0] = r.g[1] = r.g[2] = 0.0; r.cst = 1; r.aff = 1; return r; }
static sj sj_add(sj a, sj b) {
    sj r;
    r.v = a.v + b.v;
    for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i];
    r.cst = a.cst && b.cst;
    r.aff = a.aff && b.aff;
    return r;
}
static sj sj_mul(const sdf_mode *m, sj a, double s) {
    if (m->jet && s == 0.0) return sj_c(0.0);
    a.v = a.v * s;
    for (int i = 0; i < 3; ++i) a.g[i] = a.g[i] * s;
    return a;
}
static sj sj_neg(const sdf_mode *m, sj a) {
    if (!m->jet) { a.v = -a.v; return a; }
    return sj_mul(m, a, -1.0);
}
static sj sj_sub(const sdf_mode *m, sj a, sj b) {
    if (!m->jet) { a.v = a.v - b.v; return a; }
    return sj_add(a, sj_neg(m, b));
}
static sj sj_min(const sdf_mode *m, sj a, sj b) {
    sj r;
    if (!m->jet) return b.v < a.v ? b : a;                    /* std::min */
    if (a.cst && b.cst) return sj_c(b.v < a.v ? b.v : a.v);
    if (a.v + sdf_guard(a.v) < b.v - sdf_guard(b.v)) return a;
    if (b.v + sdf_guard(b.v) < a.v - sdf_guard(a.v)) return b;
    r = a.v <= b.v ? a : b;
    r.cst = 0; r.aff = 0;
    return r;
}
static sj sj_max(const sdf_mode *m, sj a, sj b) {
    if (!m->jet) return a.v < b.v ? b : a;                    /* std::max */
    return sj_neg(m, sj_min(m, sj_neg(m, a), sj_neg(m, b)));
}
static int sj_zero(sj a) { return a.v == 0.0 && a.cst; }
static sj sj_norm(const sdf_mode *m, sj a, sj b, sj c) {
    sj r, q[3];
    if (!m->jet) return sj_c(sdf_hypot3(a.v, b.v, c.v));
    if (sj_zero(b) && sj_zero(c)) return sj_max(m, a, sj_neg(m, a));
    if (sj_zero(a) && sj_zero(c)) return sj_max(m, b, sj_neg(m, b));
    if (sj_zero(a) && sj_zero(b)) return sj_max(m, c, sj_neg(m, c));
    q[0] = a; q[1] = b; q[2] = c;
    r = sj_c(sdf_hypot3(a.v, b.v, c.v));
    if (r.v > 0.0)
        for (int k = 0; k < 3; ++k) {
            double ratio = q[k].v / r.v;
            for (int i = 0; i < 3; ++i) r.g[i] = r.g[i] + q[k].g[i] * ratio;
        }
    r.cst = a.cst && b.cst && c.cst;
    if (a.aff && b.aff && c.aff) {
        int z = 1;
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) z = z && q[k].g[i] == 0.0;
        r.cst = z;
    }
    r.aff = r.cst;
    return r;
}
static sj sj_smooth_union(const sdf_mode *m, sj a, sj b, double k) {
    double d = b.v - a.v, h;
    sj r;
    if (m->jet) {
        double err = sdf_guard(a.v) + sdf_guard(b.v);
        if (d - err >= k) return a;
        if (d + err <= -k) return b;
    }
    h = 0.5 + 0.5 * d / k;
    if (h < 0.0) h = 0.0;
    else if (h > 1.0) h = 1.0;
    r = sj_c(b.v + (a.v - b.v) * h - k * h * (1.0 - h));
    for (int i = 0; i < 3; ++i) r.g[i] = h * a.g[i] + (1.0 - h) * b.g[i];
    r.cst = a.cst && b.cst;
    r.aff = r.cst;
    return r;
}
static sj sj_subtract(const sdf_mode *m, sj a, sj b, double k) {
    return sj_neg(m, sj_smooth_union(m, a, sj_neg(m, b), k));
}
static sj sj_abs(const sdf_mode *m, sj a) { return sj_max(m, a, sj_neg(m, a)); }
static void sj_shift(sj out[3], const sj *p, double x, double y, double z) {
    out[0] = sj_add(p[0], sj_c(x));
    out[1] = sj_add(p[1], sj_c(y));
    out[2] = sj_add(p[2], sj_c(z));
}
static void sj_rotate(const sdf_mode *m, sj out[3], const sj *p, const float r[9]) {
    for (int i = 0; i < 3; ++i)
        out[i] = sj_add(sj_add(sj_mul(m, p[0], (double)r[3 * i]), sj_mul(m, p[1], (double)r[3 * i + 1])),
                        sj_mul(m, p[2], (double)r[3 * i + 2]));
}
static sj sj_sphere(const sdf_mode *m, const sj *p, double r) {
    return sj_sub(m, sj_norm(m, p[0], p[1], p[2]), sj_c(r));
}
static sj sj_cone(const sdf_mode *m, const sj *p) {
    return sj_mul(m, sj_add(sj_norm(m, p[0], p[1], sj_c(0.0)), p[2]), (double)(3.1415926536f / 3.f));
}
static sj sj_cylinder(const sdf_mode *m, const sj *p, double h, double r) {
    sj x = sj_sub(m, sj_norm(m, p[0], p[2], sj_c(0.0)), sj_c(h));
    sj y = sj_sub(m, sj_abs(m, p[1]), sj_c(r));
    return sj_add(sj_min(m, sj_max(m, x, y), sj_c(0.0)),
                  sj_norm(m, sj_max(m, x, sj_c(0.0)), sj_max(m, y, sj_c(0.0)), sj_c(0.0)));
}
static sj sj_prism(const sdf_mode *m, const sj *p) {
    sj zside = sj_sub(m, sj_abs(m, p[2]), sj_c((double)0.3f));
    sj slant = sj_add(sj_mul(m, sj_abs(m, p[0]), (double)0.866025f), sj_mul(m, p[1], 0.5));
    return sj_max(m, zside, sj_sub(m, sj_max(m, slant, sj_neg(m, p[1])), sj_c((double)(1.2f * 0.5f))));
}
/* knobGeometry (SdfFunctionsQueries.cpp:245-267) */
static sj sj_knob(const sdf_mode *m, const sj *p0, int variant) {
    const double size = (double)0.8f;
    sj p[3], t[3], u[3], b[3], outer, cutout, etch, d, inner, base, cone_a, cone_b;
    for (int i = 0; i < 3; ++i) p[i] = sj_mul(m, p0[i], (double)(float)(1.0 / size));
    if (variant == ORC_SDF_KNOB_INNER) return sj_mul(m, sj_sphere(m, p, 0.75), size);
    outer = sj_sphere(m, p, 1.0);
    sj_shift(t, p, 0.0, -(double)0.5f, -(double)0.5f);
    cutout = sj_sphere(m, t, (double)0.7f);
    sj_shift(t, p, 0.0, -(double)0.2f, -(double)0.2f);
    sj_rotate(m, u, t, m->rm45);
    etch = sj_sub(m, sj_norm(m, sj_sub(m, sj_norm(m, u[0], u[2], sj_c(0.0)), sj_c((double)1.0f)), u[1], sj_c(0.0)),
                  sj_c((double)0.05f));
    d = sj_subtract(m, cutout, outer, (double)0.1f);
    inner = sj_sphere(m, p, 0.75);
    d = variant == ORC_SDF_KNOB_OUTER ? sj_max(m, d, sj_neg(m, inner)) : sj_min(m, d, inner);
    d = sj_max(m, sj_neg(m, etch), d);
    sj_shift(b, p, 0.0, (double)0.775f, 0.0);
    sj_shift(t, b, 0.0, (double)0.9f, 0.0);
    sj_rotate(m, u, t, m->rm90);
    cone_a = sj_cone(m, u);
    sj_shift(t, b, 0.0, -(double)0.9f, 0.0);
    sj_rotate(m, u, t, m->rp90);
    cone_b = sj_cone(m, u);
    base = sj_smooth_union(m, cone_a, cone_b, (double)0.02f);
    base = sj_mul(m, sj_max(m, base, sj_cylinder(m, b, (double)1.1f, (double)0.25f)), (double)0.7f);
    base = sj_max(m, sj_neg(m, sj_cylinder(m, b, (double)0.6f, (double)0.3f)), base);
    sj_shift(t, b, 0.0, 0.0, -1.0);
    sj_rotate(m, u, t, m->rp90);
    base = sj_max(m, sj_neg(m, sj_prism(m, u)), base);
    return sj_mul(m, sj_min(m, sj_subtract(m, outer, base, (double)0.1f), d), size);
}
static sj sj_slab(const sdf_mode *m, const sj *p) {
    sj x = sj_sub(m, sj_abs(m, p[0]), sj_c(0.6));
    sj y = sj_sub(m, sj_abs(m, p[1]), sj_c(0.01));
    sj z = sj_sub(m, sj_abs(m, p[2]), sj_c(0.6));
    return sj_max(m, sj_max(m, x, y), z);
}
static sj sj_two_slabs(const sdf_mode *m, const sj *p) {
    sj a[3], b[3];
    sj_shift(a, p, -0.18, -0.02, 0.0);
    sj_shift(b, p, 0.18, 0.02, 0.0);
    return sj_min(m, sj_slab(m, a), sj_slab(m, b));
}
/* Mat4f::rotAxis((1, 0, 0), angle) rows 0-2 (Mat4f.cpp:132-148), in float. */
static void sdf_rot_x(float deg, float r[9]) {
    const float x = 1.0f, y = 0.0f, z = 0.0f;
    float a = deg * (3.1415926536f / 180.0f), s = sinf(a), c = cosf(a), c1 = 1.0f - c;
    r[0] = c + x * x * c1;     r[1] = x * y * c1 - z * s; r[2] = x * z * c1 + y * s;
    r[3] = y * x * c1 + z * s; r[4] = c + y * y * c1;     r[5] = y * z * c1 - x * s;
    r[6] = z * x * c1 - y * s; r[7] = z * y * c1 + x * s; r[8] = c + z * z * c1;
}
static void sdf_mode_init(sdf_mode *m, int jet) {
    m->jet = jet;
    sdf_rot_x(90.0f, m->rp90);
    sdf_rot_x(-90.0f, m->rm90);
    sdf_rot_x(-45.0f, m->rm45);
}
double orc_sdf_jet(int fn, const double p[3], double g[3]) {
    sdf_mode m;
    sj q[3], r, s[3];
    if (fn == ORC_SDF_ROUNDED_U_GROOVE) {
        /* roundedUJet (:323-327) */
        double qy = 0.0 < p[1] ? 0.0 : p[1], rr = sqrt(p[0] * p[0] + qy * qy);
        g[0] = rr > 0.0 ? -p[0] / rr : 0.0;
        g[1] = rr > 0.0 ? -qy / rr : 0.0;
        g[2] = rr > 0.0 ? -0.0 / rr : 0.0;
        return 1.0 - rr;
    }
    sdf_mode_init(&m, 1);
    for (int i = 0; i < 3; ++i) {
        q[i] = sj_c(p[i]);
        q[i].g[i] = 1.0;
        q[i].cst = 0;
    }
    if (fn == ORC_SDF_TWO_SLABS) r = sj_two_slabs(&m, q);
    else if (fn == ORC_SDF_THREE_SLABS) {
        sj_shift(s, q, -0.18, 0.06, 0.0);
        r = sj_min(&m, sj_two_slabs(&m, q), sj_slab(&m, s));
    } else if (fn == ORC_SDF_TWO_SPHERES) {
        s[0] = q[0];
        s[1] = sj_sub(&m, sj_abs(&m, q[1]), sj_c(10.0));
        s[2] = q[2];
        r = sj_sphere(&m, s, 9.5);
    } else r = sj_knob(&m, q, ORC_SDF_KNOB);
    for (int i = 0; i < 3; ++i) g[i] = r.g[i];
    return r.v;
}
float orc_sdf_eval(int fn, const float p[3]) {
    sdf_mode m;
    sj q[3];
    if (fn == ORC_SDF_PLANE) return p[1];
    sdf_mode_init(&m, 0);
    for (int i = 0; i < 3; ++i) q[i] = sj_c((double)p[i]);
    return (float)sj_knob(&m, q, fn).v;
}
static int sdf_has_queries(int fn) {
    return fn == ORC_SDF_KNOB || fn == ORC_SDF_ROUNDED_U_GROOVE || fn == ORC_SDF_TWO_SPHERES ||
           fn == ORC_SDF_TWO_SLABS || fn == ORC_SDF_THREE_SLABS;
}
/* ProceduralMean::mean without queries (GPFunctions.cpp:282-289) */
static double sdf_mean_plain(const orc_medium_desc *d, const double a[3]) {
    const orc_f32 *t = d->sdf_inverse_transform;
    float p[3] = {(float)a[0], (float)a[1], (float)a[2]}, l[3], v;
    for (int i = 0; i < 3; ++i) l[i] = t[4 * i] * p[0] + t[4 * i + 1] * p[1] + t[4 * i + 2] * p[2] + t[4 * i + 3];
    v = orc_sdf_eval(d->sdf_function, l);
    v *= d->sdf_scale;
    v = v + d->sdf_offset;
    return (double)(d->sdf_min > v ? d->sdf_min : v);
}
double orc_sdf_mean(const orc_medium_desc *d, const double a[3], double g[3]) {
    const orc_f32 *t = d->sdf_inverse_transform;
    if (sdf_has_queries(d->sdf_function)) {
        /* ProceduralMean::evaluateJet (ProceduralMeanSdf.cpp:32-38); the 3x3 product in the
         * order Eigen evaluates it for doubles with SSE2 (rows 0-1 left to right, row 2 as
         * x + (y + z)); the gradient's rows left to right with the scale folded in first. */
        double w[9], l[3], q[3], value, s = (double)d->sdf_scale;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) w[3 * i + j] = (double)t[4 * i + j];
        l[0] = ((w[0] * a[0] + w[1] * a[1]) + w[2] * a[2]) + (double)t[3];
        l[1] = ((w[3] * a[0] + w[4] * a[1]) + w[5] * a[2]) + (double)t[7];
        l[2] = (w[6] * a[0] + (w[7] * a[1] + w[8] * a[2])) + (double)t[11];
        value = orc_sdf_jet(d->sdf_function, l, q) * s + (double)d->sdf_offset;
        if (value < (double)d->sdf_min || d->sdf_scale == 0.0f) {
            if (g) g[0] = g[1] = g[2] = 0.0;
            return value < (double)d->sdf_min ? (double)d->sdf_min : value;
        }
        if (g)
            for (int i = 0; i < 3; ++i) g[i] = ((s * w[i]) * q[0] + (s * w[3 + i]) * q[1]) + (s * w[6 + i]) * q[2];
        return value;
    }
    {
        /* dmean_da (:291-298): forward differences with eps = double(0.001f) */
        const double eps = (double)0.001f;
        double v0 = sdf_mean_plain(d, a);
        if (g)
            for (int i = 0; i < 3; ++i) {
                double b[3] = {a[0] + 0.0, a[1] + 0.0, a[2] + 0.0};
                b[i] = a[i] + eps;
                g[i] = (sdf_mean_plain(d, b) - v0) / eps;
            }
        return v0;
    }
}

/* ---- realization (SparseConvolutionNoise.cpp) ------------------------ */
static int ilog_level(const orc_field *f, float s) { return (int)floorf(logf(s) / logf(f->base)); }

/* kernelScaleLevelRatio (:101-111) */
static v4f level_ratio(const orc_field *f, v3f p) {
    float lat = cov_lateral(f, p);
    float lo = floorf(logf(lat) / logf(f->base));
    float hi = lo + 1;
    float slo = powf(f->base, lo);
    float shi = powf(f->base, hi);
    float fr = (lat - slo) / (shi - slo);
    float rlo = (float)((1.0 - fr) / sqrt(1.0 - 2.0 * fr + 2.0 * fr * fr));
    float rhi = (float)(fr / sqrt(1.0 - 2.0 * fr + 2.0 * fr * fr));
    return V4(slo, shi, rlo, rhi);
}

/* cell3D (:432-452) */
static v4f cell3d(const orc_field *f, v3f pw, uint32_t ci, uint32_t cj, uint32_t ck, v3f p,
                  uint32_t seed, float N, float R, float kss, int *pairs) {
    pcg_t g;
    pcg_set(&g, (uint64_t)(uint32_t)(orc_xxhash32_4(ck, cj, ci, seed) + 1u));
    uint32_t n = (uint32_t)N;
    v4f sum = V4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t k = 0; k < n; ++k) {
        v3f pi = pcg_3d(&g);
        float w = pcg_1d(&g) < 0.5f ? -1.f : 1.f;
        v3f to = vsub(p, pi);
        if (vlen2(to) < 1.0f) {
            sum = v4add(sum, v4scale(w, splat3d(f, vscale(R, p), vscale(R, pi), 0, f->d.iso3d != 0, kss, pw)));
        }
    }
    if (pairs) *pairs += (int)n;
    return sum;
}
/* noise3D (:419-430) */
static v4f noise3d(const orc_field *f, v3f pw, v3f p, uint32_t seed, float N, float R, float kss) {
    v3f pg = vdivs(p, R);
    v3f fl = V3(floorf(pg.x), floorf(pg.y), floorf(pg.z));
    v3f fr = vsub(pg, fl);
    int ix = (int)fl.x, iy = (int)fl.y, iz = (int)fl.z;
    v4f sum = V4(0.f, 0.f, 0.f, 0.f);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
                sum = v4add(sum, cell3d(f, pw, (uint32_t)(ix + dx), (uint32_t)(iy + dy), (uint32_t)(iz + dz),
                                        vsub(fr, V3((float)dx, (float)dy, (float)dz)), seed, N, R, kss, NULL));
    return sum;
}
/* evaluateNoise3DNormalized (:308-317) */
static v4f noise3d_normalized(const real_t *r, v3f p, float N, float R, float kss, int cond) {
    const orc_field *f = r->f;
    int addl = ilog_level(f, kss);
    v4f nz = noise3d(f, p, p, r->seed + (uint32_t)addl, N, R, kss);
    float nf = sqrtf(cov_var3d(f, p, N, R, 0, kss));
    nz = v4divs(nz, nf);
    if (f->activate && cond) {
        v4f a = v4scale(r->c.value_scale, splat3d(f, p, r->c.origin, 1, 0, kss, p));
        v4f b = splat3d_grad(f, p, r->c.origin, r->c.grad_scale, 1, 0, kss, p);
        nz = v4add(nz, v4add(a, b));
    }
    return nz;
}
/* evaluateNoise3DIsotropicRayNormalized (:345-377): the point in the isotropic space,
 * rotated into the frame of the flight's isotropic direction, noise3D there with the
 * identity-space kernel radius, the gradient rotated and scaled back to world space. */
static v4f noise3d_iso_ray_normalized(const real_t *r, v3f p, v3f dir, float N, float R, float kss, int cond) {
    const orc_field *f = r->f;
    v3f dir_iso = vnorm(w2l(f, dir, 1.0f));
    frame_t F = make_frame(dir_iso);
    v3f piso = w2l(f, p, kss);
    v3f pir = to_local(&F, piso);
    int addl = ilog_level(f, kss);
    v4f nir = noise3d(f, p, pir, r->seed + (uint32_t)addl, N, R, 1.0f);
    v3f gw = grad_l2w(f, to_global(&F, V3(nir.y, nir.z, nir.w)), kss);
    v4f nw = V4(nir.x, gw.x, gw.y, gw.z);
    float nf = sqrtf(cov_var3d(f, p, N, R, 1, 1.0f));
    nw = v4divs(nw, nf);
    if (f->activate && cond) {
        v3f oir = to_local(&F, w2l(f, r->c.origin, kss));
        v4f a = v4scale(r->c.value_scale, splat3d(f, pir, oir, 1, 1, 1.0f, p));
        v4f b = splat3d_grad(f, pir, oir, r->c.grad_scale, 1, 1, 1.0f, p);
        v4f d = v4add(a, b);
        v3f gd = grad_l2w(f, to_global(&F, V3(d.y, d.z, d.w)), kss);
        nw = v4add(nw, V4(d.x, gd.x, gd.y, gd.z));
    }
    return nw;
}
/* evaluateNoise3D (:142-165): world space, or the isotropic ray space of the flight */
static v4f eval_noise3d(const real_t *r, v3f p, v3f dir, int cond) {
    const orc_field *f = r->f;
    float N = f->d.impulse_density;
    if (f->d.iso3d) {
        float R = cov_kernel_radius(f, 1, 1.0f);
        if (!f->multires) return noise3d_iso_ray_normalized(r, p, dir, N, R, 1.0f, cond);
        v4f info = level_ratio(f, p);
        v4f lo = noise3d_iso_ray_normalized(r, p, dir, N, R, info.x, cond);
        v4f hi = noise3d_iso_ray_normalized(r, p, dir, N, R, info.y, cond);
        return v4add(v4scale(info.z, lo), v4scale(info.w, hi));
    }
    if (!f->multires) {
        float kss = cov_world_sampling_scale(f);
        return noise3d_normalized(r, p, N, cov_kernel_radius(f, 0, 1.0f), kss, cond);
    }
    v4f info = level_ratio(f, p);
    v4f lo = noise3d_normalized(r, p, N, cov_kernel_radius(f, 0, info.x), info.x, cond);
    v4f hi = noise3d_normalized(r, p, N, cov_kernel_radius(f, 0, info.y), info.y, cond);
    return v4add(v4scale(info.z, lo), v4scale(info.w, hi));
}

/* cell1D / noise1D (:454-485) */
static v2f cell1d(const orc_field *f, v3f pw, uint32_t i, float t, uint32_t seed, float N, float R) {
    pcg_t g;
    pcg_set(&g, (uint64_t)(uint32_t)(orc_xxhash32_2(i, seed) + 1u));
    uint32_t n = (uint32_t)N;
    v2f sum = {0.f, 0.f};
    for (uint32_t k = 0; k < n; ++k) {
        float ti = pcg_1d(&g);
        float w = pcg_1d(&g) < 0.5f ? -1.f : 1.f;
        float to = t - ti;
        if (to * to < 1.0f) {
            v2f s = splat1d(f, R * t, R * ti, pw);
            sum.x += w * s.x; sum.y += w * s.y;
        }
    }
    return sum;
}
static v2f noise1d(const orc_field *f, v3f pw, float t, uint32_t seed, float N, float R) {
    float tg = t / R;
    float fl = floorf(tg);
    float fr = tg - fl;
    int i = (int)fl;
    v2f sum = {0.f, 0.f};
    for (int dx = -1; dx <= 1; ++dx) {
        v2f c = cell1d(f, pw, (uint32_t)(i + dx), fr - (float)dx, seed, N, R);
        sum.x += c.x; sum.y += c.y;
    }
    return sum;
}
/* evaluateNoise1DNormalized (:379-417) */
static v4f noise1d_normalized(const real_t *r, v3f p, v3f dir, float N, float R, float kss, int cond, int low) {
    const orc_field *f = r->f;
    v3f dir_iso = vnorm(w2l(f, dir, 1.0f));
    frame_t F = make_frame(dir_iso);
    v3f piso = w2l(f, p, kss);
    v3f pir = to_local(&F, piso);
    int addl = ilog_level(f, kss);
    v2f nz = noise1d(f, p, pir.z, r->seed + (uint32_t)addl, N, R);
    float nf = sqrtf(cov_var1d(f, p, N, R));
    nz.x /= nf; nz.y /= nf;
    float gx = 0.f, gy = 0.f;
    if (f->activate && cond) {
        float osf = 1.0f;
        if (f->multires) { v4f oi = level_ratio(f, r->c.o1); osf = low ? oi.z : oi.w; }
        v3f oiso = w2l(f, r->c.o1, kss);
        v3f oir = to_local(&F, oiso);
        v2f kv = cov1d(f, pir.z, oir.z, p, r->c.o1);
        v2f kg = cov1d_grad(f, pir.z, oir.z, p, r->c.o1);
        float gz = kss * r->c.g1.z;
        v2f vs = {r->c.v1 * kv.x, r->c.v1 * kv.y};
        v2f gs = {gz * kg.x, gz * kg.y};
        nz.x += osf * (vs.x + gs.x);
        nz.y += osf * (vs.y + gs.y);
        if (f->d.correlation_xy) {
            gx = osf * kss * r->c.g1.x * cov1d_grad3n(f, pir.z, oir.z, p, r->c.o1);
            gy = osf * kss * r->c.g1.y * cov1d_grad3n(f, pir.z, oir.z, p, r->c.o1);
        }
    }
    return V4(nz.x, gx, gy, nz.y);
}
static float eval_value_noise1d(const real_t *r, v3f p, v3f dir, int cond) { /* :167-176 */
    const orc_field *f = r->f;
    float N = f->d.impulse_density;
    float R = cov_kernel_radius(f, 1, 1.0f);
    if (!f->multires) return noise1d_normalized(r, p, dir, N, R, 1.0f, cond, 0).x;
    v4f info = level_ratio(f, p);
    float lo = noise1d_normalized(r, p, dir, N, R, info.x, cond, 1).x;
    float hi = noise1d_normalized(r, p, dir, N, R, info.y, cond, 0).x;
    return info.z * lo + info.w * hi;
}
/* correlatedGradientMoments (:13-22) */
static void corr_moments(float distance, double *factor, double *variance) {
    const double x = 0.5 * (double)distance * (double)distance;
    *factor = 0.5 * exp(-0.5 * x) * (1.0 - x);
    *variance = x < 1.0 ? 0.5 * (-expm1(-x) + exp(-x) * x * (2.0 - x)) : 0.5 - 2.0 * (*factor) * (*factor);
}
/* seTransverseGradientMoments (:78-99) */
static int transverse_moments(const real_t *r, v3f p, v3f dir, int cond, double mean[2], double *var) {
    const orc_field *f = r->f;
    if (!f->d.one_d || !f->d.correlation_xy || f->d.ctxt != ORC_CTX_RENEWAL_PLUS || f->multires || !is_se(f))
        return 0;
    mean[0] = mean[1] = 0.0;
    *var = 0.5;
    if (cond && r->c.has_grad) {
        v3f delta = w2l(f, vsub(p, r->c.o1), 1.f);
        double q2 = 0.5 * dlen2(d_of(delta));
        double rho = exp(-0.5 * q2);
        *var = -0.5 * expm1(-q2);
        frame_t F = make_frame(vnorm(w2l(f, dir, 1.f)));
        v3f obs = to_local(&F, grad_w2l(f, r->c.target_grad, 1.f));
        mean[0] = rho * obs.x; mean[1] = rho * obs.y;
    }
    return 1;
}
/* evaluateGradientNoise1D (:178-296) */
static v3f eval_grad_noise1d(const real_t *r, v3f p, float tseg, v3f dir, int cond) {
    const orc_field *f = r->f;
    float N = f->d.impulse_density;
    float R = cov_kernel_radius(f, 1, 1.0f);
    v3f dir_iso = vnorm(w2l(f, dir, 1.0f));
    frame_t F = make_frame(dir_iso);
    double tm[2], tv;
    const double s2f = (double)sqrtf(2.f);
    if (transverse_moments(r, p, dir, cond, tm, &tv)) {
        v4f base = noise1d_normalized(r, p, dir, N, R, 1.f, cond, 0);
        double z[2];
        orc_standard_normals((uint64_t)r->normal_seed, 1, z);
        double s = sqrt(tv);
        v3f g = V3((float)(tm[0] + s * z[0]), (float)(tm[1] + s * z[1]), base.w);
        return grad_l2w(f, to_global(&F, g), 1.f);
    }
    float sx = ns_cov_splat_cov1d(f, p, p), sy = ns_cov_splat_cov1d(f, p, p);
    double xs0 = 1.0 / sx, xs1 = 1.0 / sy;
    int corr = (f->d.ctxt == ORC_CTX_RENEWAL_PLUS && f->d.correlation_xy);
    if (!f->multires) {
        v4f b4 = noise1d_normalized(r, p, dir, N, R, 1.0f, cond, 0);
        v3f base = V3(b4.y, b4.z, b4.w);
        pcg_t g; pcg_set(&g, (uint64_t)r->normal_seed);
        double a0, a1; box_muller(&g, &a0, &a1);
        a0 /= s2f; a1 /= s2f;
        v3f gir;
        if (tseg == 0) {
            a0 *= xs0; a1 *= xs1;
            gir = vadd(V3((float)a0, (float)a1, 0.f), base);
        } else if (!corr) {
            double c0, c1; box_muller(&g, &c0, &c1);
            c0 = c0 / s2f * xs0; c1 = c1 / s2f * xs1;
            gir = vadd(V3((float)c0, (float)c1, 0.f), base);
        } else {
            double fac, cv; corr_moments(tseg, &fac, &cv);
            double m0 = fac * 2.0 * a0, m1 = fac * 2.0 * a1;
            double c0, c1; box_muller(&g, &c0, &c1);
            double sc = sqrt(cv);
            gir = vadd(V3((float)(c0 * sc + m0), (float)(c1 * sc + m1), 0.f), base);
        }
        return grad_l2w(f, to_global(&F, gir), 1.0f);
    }
    /* multi-resolution (:244-295) */
    v4f info = level_ratio(f, p);
    v4f bl4 = noise1d_normalized(r, p, dir, N, R, info.x, cond, 1);
    v4f bh4 = noise1d_normalized(r, p, dir, N, R, info.y, cond, 0);
    v3f bl = V3(bl4.y, bl4.z, bl4.w), bh = V3(bh4.y, bh4.z, bh4.w);
    pcg_t g; pcg_set(&g, (uint64_t)r->normal_seed);
    double l0, l1, h0, h1;
    box_muller(&g, &l0, &l1); l0 /= s2f; l1 /= s2f;
    box_muller(&g, &h0, &h1); h0 /= s2f; h1 /= s2f;
    v3f gl, gh;
    if (tseg == 0) {
        gl = vadd(V3((float)(l0 * xs0), (float)(l1 * xs1), 0.f), bl);
        gh = vadd(V3((float)(h0 * xs0), (float)(h1 * xs1), 0.f), bh);
    } else if (!corr) {
        double a, b, c, e;
        box_muller(&g, &a, &b); a = a / s2f * xs0; b = b / s2f * xs1;
        box_muller(&g, &c, &e); c = c / s2f * xs0; e = e / s2f * xs1;
        gl = vadd(V3((float)a, (float)b, 0.f), bl);
        gh = vadd(V3((float)c, (float)e, 0.f), bh);
    } else {
        double fac, cv; corr_moments(tseg, &fac, &cv);
        double a, b, c, e; double sc = sqrt(cv);
        box_muller(&g, &a, &b); box_muller(&g, &c, &e);
        gl = vadd(V3((float)(a * sc + fac * 2.0 * l0), (float)(b * sc + fac * 2.0 * l1), 0.f), bl);
        gh = vadd(V3((float)(c * sc + fac * 2.0 * h0), (float)(e * sc + fac * 2.0 * h1), 0.f), bh);
    }
    frame_t Fl = make_frame(vnorm(w2l(f, dir, info.x)));
    frame_t Fh = make_frame(vnorm(w2l(f, dir, info.y)));
    v3f wl = grad_l2w(f, to_global(&Fl, gl), info.x);
    v3f wh = grad_l2w(f, to_global(&Fh, gh), info.y);
    return vadd(vscale(info.z, wl), vscale(info.w, wh));
}

/* evaluateValue / evaluateGradient (:114-140) */
static float eval_value(const real_t *r, v3f p, v3f dir) {
    const orc_field *f = r->f;
    float amp = cov_amplitude(f, p);
    float nv = f->d.one_d ? eval_value_noise1d(r, p, dir, 1) : eval_noise3d(r, p, dir, 1).x;
    double m = mean_value(f, d_of(p));
    return (float)((double)(amp * nv) + m);
}
static v3f eval_gradient(const real_t *r, v3f p, float t, v3f dir) {
    const orc_field *f = r->f;
    float amp = cov_amplitude(f, p);
    v3f ng;
    if (f->d.one_d) ng = eval_grad_noise1d(r, p, t, dir, 1);
    else { v4f n4 = eval_noise3d(r, p, dir, 1); ng = V3(n4.y, n4.z, n4.w); }
    v3f mg = f_of(mean_grad(f, d_of(p)));
    return vadd(vscale(amp, ng), mg);
}

/* A marched value, optionally perturbed by +-perturb with a hashed sign: the test
 * hook that decides whether a result is stable under FP32-sized evaluation error. */
static float march_value(real_t *r, v3f p, v3f dir) {
    float v = eval_value(r, p, dir);
    if (r->perturb != 0.f) {
        r->pstate = orc_xxhash32_1(r->pstate);
        v += (r->pstate & 1u) ? r->perturb : -r->perturb;
    }
    return v;
}

/* conditioning3D (:498-574), world space */
static void conditioning3d(real_t *r, v3f p, v3f dir, float tv, v3f tg) {
    const orc_field *f = r->f;
    r->c.value_scale = 0.f; r->c.grad_scale = V3(0.f, 0.f, 0.f); r->c.origin = p;
    v4f ksi = V4(0, 0, 0, 0);
    if (f->multires) ksi = level_ratio(f, p);
    float amp = 0.f;
    if (f->d.ctxt == ORC_CTX_RENEWAL || f->d.ctxt == ORC_CTX_RENEWAL_PLUS) {
        amp = cov_amplitude(f, p);
        if (amp == 0) return;
        float mean = (float)mean_value(f, d_of(p));
        float cur = eval_noise3d(r, p, dir, 0).x;
        r->c.value_scale = (tv - mean) / amp - cur;
        if (f->multires) r->c.value_scale /= ksi.z + ksi.w;
    }
    if (f->d.ctxt == ORC_CTX_RENEWAL_PLUS) {
        v3f mg = f_of(mean_grad(f, d_of(p)));
        v4f n4 = eval_noise3d(r, p, dir, 0);
        v3f cur = V3(n4.y, n4.z, n4.w);
        v3f delta = vsub(vdivs(vsub(tg, mg), amp), cur);
        if (f->d.iso3d) {
            /* the isotropic (ray) space: localToWorld^T delta in the flight's frame (:539-556) */
            v3f gsi = grad_w2l(f, delta, 1.0f);
            frame_t F = make_frame(vnorm(w2l(f, dir, 1.0f)));
            gsi = to_local(&F, gsi);
            m3f inv = one_over_second(f, 1);
            r->c.grad_scale = m3_mul(&inv, gsi);
            if (f->multires) {
                r->c.grad_scale = vdivs(r->c.grad_scale, ksi.z / ksi.x + ksi.w / ksi.y);
            } else {
                float s = sqrf(cov_splat_scale(f, p));
                r->c.grad_scale = V3(r->c.grad_scale.x * s, r->c.grad_scale.y * s, r->c.grad_scale.z * s);
            }
        } else {
            m3f inv = one_over_second(f, 0);
            r->c.grad_scale = m3_mul(&inv, delta);
            if (f->multires) {
                float den = ksi.z / (ksi.x * ksi.x) + ksi.w / (ksi.y * ksi.y);
                r->c.grad_scale = vdivs(r->c.grad_scale, den);
            } else {
                float s = sqrf(cov_lateral(f, p));
                r->c.grad_scale = V3(r->c.grad_scale.x * s, r->c.grad_scale.y * s, r->c.grad_scale.z * s);
            }
        }
    }
}
/* conditioning1D (:576-656) */
static void conditioning1d(real_t *r, v3f p, v3f dir, float tv, v3f tg) {
    const orc_field *f = r->f;
    r->c.v1 = 0.f; r->c.g1 = V3(0.f, 0.f, 0.f); r->c.o1 = p; r->c.has_grad = 0;
    v4f ksi = V4(0, 0, 0, 0);
    if (f->multires) ksi = level_ratio(f, p);
    float amp = 0.f;
    if (f->d.ctxt == ORC_CTX_RENEWAL || f->d.ctxt == ORC_CTX_RENEWAL_PLUS) {
        amp = cov_amplitude(f, p);
        if (amp == 0) return;
        float mean = (float)mean_value(f, d_of(p));
        float cur = eval_value_noise1d(r, p, dir, 0);
        r->c.v1 = (tv - mean) / amp - cur;
        if (f->multires) r->c.v1 /= sqrf(ksi.z) + sqrf(ksi.w);
    }
    if (f->d.ctxt == ORC_CTX_RENEWAL_PLUS) {
        v3f mg = f_of(mean_grad(f, d_of(p)));
        v3f cur = eval_grad_noise1d(r, p, 0.f, dir, 0);
        v3f delta = vsub(vdivs(vsub(tg, mg), amp), cur);
        v3f dir_iso = vnorm(w2l(f, dir, 1.0f));
        frame_t F = make_frame(dir_iso);
        v3f gsi = to_local(&F, grad_w2l(f, delta, 1.0f));
        float sz = ns_cov_splat_cov1d(f, p, p), sx = sz, sy = sz;
        r->c.g1 = vmul(vscale(-2.f, gsi), V3(sqrf(sx), sqrf(sy), sqrf(sz)));
        if (f->multires) r->c.g1 = vdivs(r->c.g1, sqrf(ksi.z) + sqrf(ksi.w));
        if (!f->d.correlation_xy) { r->c.g1.x = 0; r->c.g1.y = 0; }
        r->c.target_grad = vdivs(vsub(tg, mg), amp);
        r->c.has_grad = 1;
    }
}
static void conditioning(real_t *r, v3f p, v3f dir, float tv, v3f tg) { /* :488-496 */
    if (!r->f->activate) return;
    if (r->f->d.one_d) conditioning1d(r, p, dir, tv, tg);
    else conditioning3d(r, p, dir, tv, tg);
}
/* samplingScheme (:68-76) */
static int sampling_scheme(const real_t *r, v3f p) {
    const orc_field *f = r->f;
    float amp = cov_amplitude(f, p);
    if (amp < 1e-6f || (f->d.correlation_xy && r->c.has_grad && p.x == r->c.o1.x && p.y == r->c.o1.y &&
                        p.z == r->c.o1.z && is_se(f)))
        return ORC_SCHEME_UNI;
    return f->scheme1d;
}

/* ===================================================================== */
/* Public field API                                                      */
/* ===================================================================== */

orc_field *orc_field_create(const orc_medium_desc *desc) {
    orc_field *f = (orc_field *)calloc(1, sizeof(orc_field));
    f->d = *desc;
    /* SquaredExponentialCovariance::fromJson (GPFunctions.cpp:657-685) */
    f->l_conv = f->d.length_scale * sqrtf(2.f) / 2;
    f->l_aniso = vscale(f->l_conv, V3(f->d.aniso[0], f->d.aniso[1], f->d.aniso[2]));
    f->l_aniso_inv = V3(1.0f / f->l_aniso.x, 1.0f / f->l_aniso.y, 1.0f / f->l_aniso.z);
    if (isinf(f->l_aniso_inv.x) || isnan(f->l_aniso_inv.x)) f->l_aniso_inv.x = 0;
    if (isinf(f->l_aniso_inv.y) || isnan(f->l_aniso_inv.y)) f->l_aniso_inv.y = 0;
    if (isinf(f->l_aniso_inv.z) || isnan(f->l_aniso_inv.z)) f->l_aniso_inv.z = 0;
    /* SparseConvolutionNoiseRealization ctor (SparseConvolutionNoise.cpp:25-54) */
    f->activate = !f->d.single_realization && (f->d.ctxt == ORC_CTX_RENEWAL || f->d.ctxt == ORC_CTX_RENEWAL_PLUS);
    f->scheme1d = ORC_SCHEME_UNI;
    if (!f->d.single_realization && f->d.one_d) f->scheme1d = f->d.scheme;
    f->multires = !is_se(f) && f->d.ns_multires;
    f->base = 2.5f;
    if (f->d.use_aniso_mtx) {
        m3f A;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) { A.m[i][j] = f->l_conv * f->d.aniso_mtx[3 * i + j]; f->l2w[i][j] = A.m[i][j]; }
        m3f W = m3_inverse(A);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) f->w2l[i][j] = W.m[i][j];
        /* _cov_mtx_inv = W^T W */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                float s = W.m[0][i] * W.m[0][j];
                s += W.m[1][i] * W.m[1][j];
                s += W.m[2][i] * W.m[2][j];
                f->cinv[i][j] = s;
            }
        const float (*c)[3] = f->cinv;
        f->cinv_det = c[0][0] * (c[1][1] * c[2][2] - c[1][2] * c[2][1]) - c[0][1] * (c[1][0] * c[2][2] - c[1][2] * c[2][0]) +
                      c[0][2] * (c[1][0] * c[2][1] - c[1][1] * c[2][0]);
        f->full = 1;
    }
    return f;
}
void orc_field_destroy(orc_field *f) { free(f); }
float orc_kernel_radius(const orc_field *f) { return cov_kernel_radius(f, 0, 1.0f); }

/* computeSeed (:57-66) */
uint32_t orc_compute_seed(const orc_field *f, const uint32_t pss[4], uint32_t scene_seed) {
    uint32_t seed = f->d.global_seed;
    if (!f->d.single_realization) {
        uint32_t w = f->d.ctxt == ORC_CTX_GLOBAL ? 0u : pss[3];
        seed += orc_xxhash32_4(pss[0], pss[1], pss[2], w) + orc_xxhash32_1(scene_seed);
    }
    return seed;
}
static void real_init(real_t *r, const orc_field *f, const uint32_t pss[4], uint32_t scene_seed) {
    memset(r, 0, sizeof *r);
    r->f = f;
    r->seed = orc_compute_seed(f, pss, scene_seed);
    r->normal_seed = orc_xxhash32_4(pss[0], pss[1], pss[2], pss[3]) + orc_xxhash32_1(scene_seed) + 1u;
}

void orc_noise3d(const orc_field *f, const float *p, uint32_t seed, float out4[4]) {
    real_t r; memset(&r, 0, sizeof r); r.f = f; r.seed = seed;
    v4f n = eval_noise3d(&r, V3(p[0], p[1], p[2]), V3(0.f, 0.f, 1.f), 0);
    out4[0] = n.x; out4[1] = n.y; out4[2] = n.z; out4[3] = n.w;
}

typedef struct {
    const orc_field *f; const float *xyz; float *out; int64_t n0, n1; uint32_t seed;
    const orc_ray *rays; orc_flight *fl; v3f dir;
} job_t;

/* evaluateValue / evaluateGradient of the unconditioned realization `seed` at q for a
 * flight along `dir` (the direction matters only to the isotropic ray space). */
static void eval_point(const orc_field *f, uint32_t seed, const float *q, v3f dir, float *o) {
    real_t r; memset(&r, 0, sizeof r); r.f = f; r.seed = seed;
    v3f p = V3(q[0], q[1], q[2]);
    o[0] = eval_value(&r, p, dir);
    v3f g = eval_gradient(&r, p, 1.0f, dir);
    o[1] = g.x; o[2] = g.y; o[3] = g.z;
}
static void *eval_worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (int64_t i = j->n0; i < j->n1; ++i) eval_point(j->f, j->seed, j->xyz + 3 * i, j->dir, j->out + 4 * i);
    return NULL;
}
void orc_eval_points_dir(const orc_field *f, const float *xyz, int64_t n, uint32_t seed, const float dir[3],
                         float *out4, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; job_t jb[256];
    for (int t = 0; t < nthreads; ++t) {
        jb[t].f = f; jb[t].xyz = xyz; jb[t].out = out4; jb[t].seed = seed; jb[t].dir = V3(dir[0], dir[1], dir[2]);
        jb[t].n0 = n * t / nthreads; jb[t].n1 = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, eval_worker, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
void orc_eval_points(const orc_field *f, const float *xyz, int64_t n, uint32_t seed, float *out4, int nthreads) {
    const float dz[3] = {0.f, 0.f, 1.f};
    orc_eval_points_dir(f, xyz, n, seed, dz, out4, nthreads);
}

/* prepareGpisSegment (SparseConvolutionNoiseMedium.cpp:199-263) */
static void prepare(real_t *r, const orc_field *f, const orc_ray *ray) {
    real_init(r, f, ray->pss, ray->scene_seed);
    v3f p = f_of(d_of(V3(ray->origin[0], ray->origin[1], ray->origin[2])));
    v3f dir = V3(ray->dir[0], ray->dir[1], ray->dir[2]);
    if (!ray->first_scatter) {
        v3f tg = V3((float)ray->last_aniso[0], (float)ray->last_aniso[1], (float)ray->last_aniso[2]);
        conditioning(r, p, dir, (float)(double)ray->last_val, tg);
    }
}

static void trace_flight_impl(const orc_field *f, const orc_ray *ray, orc_flight *out, float perturb, uint32_t pseed);
void orc_trace_flight(const orc_field *f, const orc_ray *ray, orc_flight *out) { trace_flight_impl(f, ray, out, 0.f, 0u); }
void orc_trace_flight_perturbed(const orc_field *f, const orc_ray *ray, float eps, uint32_t pseed, orc_flight *out) {
    trace_flight_impl(f, ray, out, eps, pseed);
}
static void trace_flight_impl(const orc_field *f, const orc_ray *ray, orc_flight *out, float perturb, uint32_t pseed) {
    memset(out, 0, sizeof *out);
    int evals = 0;
    /* beginDistance (GaussianProcessMedium.cpp:236-287) */
    float near_t = ray->near_t, far_t = ray->far_t;
    if (!isfinite(far_t)) far_t = near_t + 2000;
    float maxT = far_t;
    out->value = ray->last_val;
    if (ray->bounce >= f->d.max_bounce) {
        out->outcome = ORC_FAIL; out->fail_reason = ORC_FAIL_BOUNCE_LIMIT; out->t = maxT;
        return;
    }
    if (maxT <= near_t) {
        out->outcome = ORC_EXIT; out->t = maxT; out->scheme = ORC_SCHEME_UNI;
        return;
    }
    real_t r;
    prepare(&r, f, ray);
    r.perturb = perturb;
    r.pstate = pseed;
    v3f dirf = V3(ray->dir[0], ray->dir[1], ray->dir[2]);
    v3d p = D3(ray->origin[0], ray->origin[1], ray->origin[2]);
    v3d rd = d_of(dirf);
    /* finishGpisSegment, CPU march (SparseConvolutionNoiseMedium.cpp:287-346) */
    float step_size = (far_t - near_t) / (float)f->d.min_step;
    if (f->d.step_size < step_size) step_size = f->d.step_size;
    double t = near_t;
    double f0 = march_value(&r, f_of(D3(p.x + t * rd.x, p.y + t * rd.y, p.z + t * rd.z)), dirf); ++evals;
    float minf = ray->first_scatter ? (float)fabs(f0) : INFINITY;
    int sign0 = f0 < 0 ? -1 : 1;
    double pf = f0;
    t = near_t + step_size * ray->jitter;
    int step = 0, hit = 0;
    while (t < far_t) {
        step++;
        double fc = march_value(&r, f_of(D3(p.x + t * rd.x, p.y + t * rd.y, p.z + t * rd.z)), dirf); ++evals;
        int sc = fc < 0 ? -1 : 1;
        if ((float)fabs(fc) < minf) minf = (float)fabs(fc);
        if (!ray->first_scatter && step == 1) {
            sign0 = sc;
        } else if (sc != sign0) {
            double fac = pf / (pf - fc);
            out->bracket[0] = (float)pf; out->bracket[1] = (float)fc;
            double a = t - step_size, b = t;
            double tp = a * (1.0 - fac) + b * fac, tt;
            for (;;) {
                tt = a * (1.0 - fac) + b * fac;
                double ft = march_value(&r, f_of(D3(p.x + tt * rd.x, p.y + tt * rd.y, p.z + tt * rd.z)), dirf); ++evals;
                int st = ft < 0 ? -1 : 1;
                if ((float)fabs(ft) < minf) minf = (float)fabs(ft);
                if (st == sign0) break;
                fac *= 0.9;
                if (fac <= 0.01) { tp = tt = 0; break; }
                tp = tt;
            }
            t = tp;
            hit = 1;
            break;
        }
        pf = fc;
        t += step_size;
    }
    out->steps = step;
    out->min_abs_f = minf;
    (void)0;
    if (hit) {
        out->value = 0.0f;
    } else {
        t = far_t;
        out->value = eval_value(&r, f_of(D3(p.x + t * rd.x, p.y + t * rd.y, p.z + t * rd.z)), dirf); ++evals;
    }
    out->t = t;
    /* resumeDistance (GaussianProcessMedium.cpp:289-369): gradient, rejection rules */
    v3d ro = p, rdn = dnorm(rd);
    v3f ip = f_of(D3(ro.x + rdn.x * t, ro.y + rdn.y * t, ro.z + rdn.z * t));
    v3f g = eval_gradient(&r, ip, (float)t, dirf); ++evals;
    out->grad[0] = g.x; out->grad[1] = g.y; out->grad[2] = g.z;
    out->evals = evals;
    if (hit) {
        out->scheme = sampling_scheme(&r, f_of(D3(p.x + t * rd.x, p.y + t * rd.y, p.z + t * rd.z)));
        v3d gd = d_of(g);
        double avg = (gd.x + gd.y + gd.z) * (1.0 / 3.0);
        if (!isfinite(avg)) { out->outcome = ORC_FAIL; out->fail_reason = ORC_FAIL_NONFINITE; return; }
        if (ddot(gd, d_of(dirf)) > 0) { out->outcome = ORC_FAIL; out->fail_reason = ORC_FAIL_OUTWARD; return; }
        if (dlen2(gd) < (double)0.0000001f) { out->outcome = ORC_FAIL; out->fail_reason = ORC_FAIL_ZERO_GRAD; return; }
        out->outcome = ORC_HIT;
    } else {
        out->scheme = ORC_SCHEME_UNI;
        out->outcome = ORC_EXIT;
    }
}

static void *flight_worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (int64_t i = j->n0; i < j->n1; ++i) orc_trace_flight(j->f, &j->rays[i], &j->fl[i]);
    return NULL;
}
void orc_trace_flights(const orc_field *f, const orc_ray *rays, int64_t n, orc_flight *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; job_t jb[256];
    for (int t = 0; t < nthreads; ++t) {
        jb[t].f = f; jb[t].rays = rays; jb[t].fl = out;
        jb[t].n0 = n * t / nthreads; jb[t].n1 = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, flight_worker, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

void orc_conditioned_eval(const orc_field *f, const orc_ray *ray, const float q[3], float t_segment, float out4[4]) {
    real_t r;
    prepare(&r, f, ray);
    v3f dir = V3(ray->dir[0], ray->dir[1], ray->dir[2]);
    v3f p = V3(q[0], q[1], q[2]);
    out4[0] = eval_value(&r, p, dir);
    v3f g = eval_gradient(&r, p, t_segment, dir);
    out4[1] = g.x; out4[2] = g.y; out4[3] = g.z;
}

/* seTransverseGradientMoments and samplingScheme at p for the flight's realization
 * (conditioned at the ray origin as prepare() does); returns 1 when the correlated
 * transverse law applies (SparseConvolutionNoise.cpp:68-99). */
int orc_transverse_moments(const orc_field *f, const orc_ray *ray, const float p_in[3], double mean_out[2],
                           double *var_out, int *scheme_out) {
    real_t r;
    prepare(&r, f, ray);
    v3f dir = V3(ray->dir[0], ray->dir[1], ray->dir[2]);
    v3f p = V3(p_in[0], p_in[1], p_in[2]);
    int ok = transverse_moments(&r, p, dir, 1, mean_out, var_out);
    *scheme_out = sampling_scheme(&r, p);
    return ok;
}

/* neeShared / neeGrad / neePDF (SparseConvolutionNoise.cpp:662-821), single-level SE */
void orc_nee(const orc_field *f, const orc_ray *ray, const float normal_in[3], const float p_in[3],
             float t_segment, float grad_out[3], float *pdf_out) {
    real_t r;
    prepare(&r, f, ray);
    v3f dir = V3(ray->dir[0], ray->dir[1], ray->dir[2]);
    v3f normal = V3(normal_in[0], normal_in[1], normal_in[2]);
    v3f p = V3(p_in[0], p_in[1], p_in[2]);
    float amp = cov_amplitude(f, p);
    v3f mg = f_of(mean_grad(f, d_of(p)));
    v3f dir_iso = vnorm(w2l(f, dir, 1.0f));
    frame_t F = make_frame(dir_iso);
    /* mtx_pt = amp * localToWorldInvTranspose * [t b n] */
    m3f M;
    v3f cols[3] = {F.tangent, F.bitangent, F.normal};
    float ltwit[3] = {f->l_aniso_inv.x, f->l_aniso_inv.y, f->l_aniso_inv.z};
    float sc = is_se(f) ? 1.0f : (f->multires ? 1.0f : cov_ls_max(f));
    for (int c = 0; c < 3; ++c) {
        float cv[3] = {cols[c].x, cols[c].y, cols[c].z};
        for (int rr = 0; rr < 3; ++rr) M.m[rr][c] = (amp * (ltwit[rr] / sc)) * cv[rr];
    }
    m3f Minv = m3_inverse(M);
    v4f gc4 = noise1d_normalized(&r, p, dir, f->d.impulse_density, cov_kernel_radius(f, 1, 1.0f), 1.0f, 1, 0);
    v3f gc = V3(gc4.y, gc4.z, gc4.w);
    v3f ppt = vadd(m3_mul(&M, V3(0.f, 0.f, gc.z)), mg);
    v3f xa = m3_mul(&M, V3(1.f, 0.f, 0.f)), ya = m3_mul(&M, V3(0.f, 1.f, 0.f));
    v3f cr = vcross(xa, ya);
    v3f pn = vnorm(cr);
    float stretch = vlen(cr);
    float isect = vdot(ppt, pn) / vdot(normal, pn);
    v3f g3w = vscale(isect, normal);
    v3f g3 = m3_mul(&Minv, vsub(g3w, mg));
    grad_out[0] = g3w.x; grad_out[1] = g3w.y; grad_out[2] = g3w.z;
    if (isect < 0.0) { *pdf_out = 0.f; return; }
    float sx = ns_cov_splat_cov1d(f, p, p), sy = sx;
    float pdf_area;
    double tm[2], tv;
    if (transverse_moments(&r, p, dir, 1, tm, &tv)) {
        if (!(tv > 0.)) { *pdf_out = 0.f; return; }
        double dx = g3.x - tm[0], dy = g3.y - tm[1];
        pdf_area = (float)(exp(-0.5 * (dx * dx + dy * dy) / tv) / (2.0 * M_PI * tv));
    } else if (!(f->d.ctxt == ORC_CTX_RENEWAL_PLUS && f->d.correlation_xy) || t_segment == 0.0f) {
        float smx = (float)((g3.x - gc.x) * sqrt(2.0) * sx);
        float smy = (float)((g3.y - gc.y) * sqrt(2.0) * sy);
        pdf_area = (float)(exp(-smx * smx / 2.0) * exp(-smy * smy / 2.0) / (2.0 * M_PI) * 2.0 * sx * sy);
    } else {
        double a0, a1;
        pcg_t g; pcg_set(&g, (uint64_t)r.normal_seed);
        box_muller(&g, &a0, &a1);
        double s2f = (double)sqrtf(2.f);
        a0 /= s2f; a1 /= s2f;
        double fac, cv; corr_moments(t_segment, &fac, &cv);
        double m0 = fac * 2.0 * a0, m1 = fac * 2.0 * a1;
        float smx = (float)((g3.x - gc.x - m0) / sqrt(cv));
        float smy = (float)((g3.y - gc.y - m1) / sqrt(cv));
        pdf_area = (float)(exp(-smx * smx / 2.0) * exp(-smy * smy / 2.0) / (2.0 * M_PI) / cv);
    }
    float cosl = fabsf(vdot(normal, pn));
    float pdf_n = pdf_area * (isect * isect) / cosl / stretch;
    *pdf_out = (float)(pdf_n / (4.0 * vdot(normal, vscale(-1.f, dir))));
}

/* ---- per-ray length scale -------------------------------------------- */
typedef struct {
    const orc_medium_desc *d; const orc_ray *rays; const float *ls; orc_flight *fl; int64_t n0, n1;
} ls_job_t;
static void *ls_worker(void *arg) {
    ls_job_t *j = (ls_job_t *)arg;
    for (int64_t i = j->n0; i < j->n1; ++i) {
        orc_medium_desc d = *j->d;
        if (j->ls) d.length_scale = j->ls[i];
        orc_field *f = orc_field_create(&d);
        orc_trace_flight(f, &j->rays[i], &j->fl[i]);
        orc_field_destroy(f);
    }
    return NULL;
}
void orc_trace_flights_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, int64_t n,
                          orc_flight *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; ls_job_t jb[256];
    for (int t = 0; t < nthreads; ++t) {
        jb[t].d = desc; jb[t].rays = rays; jb[t].ls = ls; jb[t].fl = out;
        jb[t].n0 = n * t / nthreads; jb[t].n1 = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, ls_worker, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
void orc_nee_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, const float *normals,
                const float *points, const float *t_segment, int64_t n, float *grads, float *pdfs) {
    for (int64_t i = 0; i < n; ++i) {
        orc_medium_desc d = *desc;
        if (ls) d.length_scale = ls[i];
        orc_field *f = orc_field_create(&d);
        orc_nee(f, &rays[i], normals + 3 * i, points + 3 * i, t_segment[i], grads + 3 * i, pdfs + i);
        orc_field_destroy(f);
    }
}
void orc_conditioned_eval_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls,
                             const float *q, const float *t_segment, int64_t n, float *out4) {
    for (int64_t i = 0; i < n; ++i) {
        orc_medium_desc d = *desc;
        if (ls) d.length_scale = ls[i];
        orc_field *f = orc_field_create(&d);
        orc_conditioned_eval(f, &rays[i], q + 3 * i, t_segment[i], out4 + 4 * i);
        orc_field_destroy(f);
    }
}

/* ===================================================================== */
/* 1D paths: flights chained through the hit-time phase                  */
/* ===================================================================== */

/* Exit distance of o + t d from the box [-b, b]^3 (the harness's scene bound). */
static float box_exit_f(v3f o, v3f d, float b) {
    float tx = d.x != 0.f ? fmaxf((b - o.x) / d.x, (-b - o.x) / d.x) : 3.4e38f;
    float ty = d.y != 0.f ? fmaxf((b - o.y) / d.y, (-b - o.y) / d.y) : 3.4e38f;
    float tz = d.z != 0.f ? fmaxf((b - o.z) / d.z, (-b - o.z) / d.z) : 3.4e38f;
    return fmaxf(fminf(tx, fminf(ty, tz)), 0.f);
}
static void path_stream(pcg_t *g, uint32_t path, uint32_t b, uint32_t seed, uint32_t stream) {
    pcg_set(g, (uint64_t)(uint32_t)(orc_xxhash32_4(path, b, seed, stream) + 1u));
}
static void trace_path1d(const orc_field *f, const orc_ray *ray0, uint32_t path, int max_bounces, int phase,
                         uint32_t bseed, float box, const float *light, orc_path *out) {
    const float INV_PI_F = 1.0f / PI_F, TWO_PI_F = PI_F * 2.0f; /* Angle.hpp:8-14 */
    memset(out, 0, sizeof *out);
    orc_ray ray = *ray0;
    const uint32_t w0 = ray0->pss[3];
    for (int b = 0;; ++b) {
        ray.pss[3] = w0 + (uint32_t)b;
        ray.bounce = (int32_t)(w0 + (uint32_t)b);
        orc_flight fl;
        orc_trace_flight(f, &ray, &fl);
        out->flights++;
        out->evals += fl.evals;
        if (b == 0) {
            out->t0 = (float)fl.t;
            out->grad0[0] = fl.grad[0]; out->grad0[1] = fl.grad[1]; out->grad0[2] = fl.grad[2];
            out->outcome0 = fl.outcome; out->fail0 = fl.fail_reason;
        }
        out->t_last = (float)fl.t;
        out->outcome_last = fl.outcome; out->fail_last = fl.fail_reason;
        for (int a = 0; a < 3; ++a) { out->pos_last[a] = ray.origin[a]; out->dir_last[a] = ray.dir[a]; }
        if (fl.outcome != ORC_HIT) break;
        out->hits++;
        /* finishDistance: sample.p = ray.pos() + sample.t * ray.dir() in float */
        const float t = (float)fl.t;
        const v3f o = V3(ray.origin[0], ray.origin[1], ray.origin[2]);
        const v3f d = V3(ray.dir[0], ray.dir[1], ray.dir[2]);
        const v3f p = vadd(o, vscale(t, d));
        const v3f g = V3(fl.grad[0], fl.grad[1], fl.grad[2]);
        /* setEventIsectInfo: TangentFrame(Vec3f(aniso.normalized())), wi = toLocal(-w).normalized() */
        const frame_t F = make_frame(f_of(dnorm(d_of(g))));
        const v3f wi = vnorm(to_local(&F, V3(-d.x, -d.y, -d.z)));
        if (light) {
            v3f dl = V3(light[0] - p.x, light[1] - p.y, light[2] - p.z);
            const float rsq = vlen2(dl);
            const float dist = sqrtf(rsq);
            dl = vdivs(dl, dist);
            float fv = 0.f;
            v3f sg = g;
            const v3f wo = vnorm(to_local(&F, dl));
            if (phase == ORC_PHASE_MIRROR) {
                /* specular with the UNI scheme: no light sample (TraceBase.cpp:806-807) */
                if (fl.scheme != ORC_SCHEME_UNI) {
                    /* MirrorBsdf::eval / evalGrad: the half-vector normal in world space */
                    const v3f nl = vscale(0.5f, vadd(wi, wo));
                    const v3f h = vnorm(to_global(&F, nl));
                    const float hh[3] = {h.x, h.y, h.z}, pp[3] = {p.x, p.y, p.z};
                    float gg[3], pdf;
                    orc_nee(f, &ray, hh, pp, t, gg, &pdf);
                    fv = pdf;
                    sg = V3(gg[0], gg[1], gg[2]);
                }
            } else if (wi.z > 0.0f && wo.z > 0.0f) {
                fv = INV_PI_F * wo.z; /* LambertBsdf::eval */
            }
            if (fv != 0.0f) {
                orc_ray sh = ray;
                sh.origin[0] = p.x; sh.origin[1] = p.y; sh.origin[2] = p.z;
                sh.dir[0] = dl.x; sh.dir[1] = dl.y; sh.dir[2] = dl.z;
                sh.near_t = 0.0f;
                sh.far_t = fminf(dist, box_exit_f(p, dl, box));
                sh.first_scatter = 0;
                sh.last_val = 0.0f;
                sh.last_aniso[0] = sg.x; sh.last_aniso[1] = sg.y; sh.last_aniso[2] = sg.z;
                sh.bounce = 0; /* transmittance has no bounce limit */
                pcg_t gs;
                path_stream(&gs, path, (uint32_t)b, bseed, 1u);
                sh.jitter = pcg_1d(&gs);
                orc_flight sf;
                orc_trace_flight(f, &sh, &sf);
                out->shadows++;
                out->evals += sf.evals;
                if (sf.outcome == ORC_EXIT) { /* any crossing occludes */
                    out->visible++;
                    out->nee += fv / rsq;
                }
            }
        }
        if (b == max_bounces) break;
        /* phase sample of the next flight */
        pcg_t gb;
        path_stream(&gb, path, (uint32_t)(b + 1), bseed, 0u);
        const float u1 = pcg_1d(&gb), u2 = pcg_1d(&gb), jit = pcg_1d(&gb);
        v3f wl;
        if (phase == ORC_PHASE_MIRROR) {
            wl = V3(-wi.x, -wi.y, wi.z); /* MirrorBsdf::sample */
        } else {
            if (wi.z <= 0.0f) break; /* LambertBsdf::sample refuses */
            const float phi = u1 * TWO_PI_F, r = sqrtf(u2); /* SampleWarp::cosineHemisphere */
            wl = V3(cosf(phi) * r, sinf(phi) * r, sqrtf(fmaxf(1.0f - u2, 0.0f)));
        }
        const v3f w = vnorm(to_global(&F, wl));
        ray.origin[0] = p.x; ray.origin[1] = p.y; ray.origin[2] = p.z;
        ray.dir[0] = w.x; ray.dir[1] = w.y; ray.dir[2] = w.z;
        ray.near_t = 0.0f;
        ray.far_t = box_exit_f(p, w, box);
        ray.jitter = jit;
        ray.first_scatter = 0;
        ray.last_val = 0.0f;
        ray.last_aniso[0] = g.x; ray.last_aniso[1] = g.y; ray.last_aniso[2] = g.z;
    }
}

typedef struct {
    const orc_medium_desc *d; const orc_ray *rays; const float *ls; orc_path *out; int64_t n0, n1;
    int max_bounces, phase; uint32_t bseed; float box; const float *light;
} path_job_t;
static void *path_worker(void *arg) {
    path_job_t *j = (path_job_t *)arg;
    orc_field *shared = j->ls ? NULL : orc_field_create(j->d);
    for (int64_t i = j->n0; i < j->n1; ++i) {
        orc_field *f = shared;
        if (!f) {
            orc_medium_desc d = *j->d;
            d.length_scale = j->ls[i];
            f = orc_field_create(&d);
        }
        trace_path1d(f, &j->rays[i], (uint32_t)i, j->max_bounces, j->phase, j->bseed, j->box, j->light, &j->out[i]);
        if (!shared) orc_field_destroy(f);
    }
    if (shared) orc_field_destroy(shared);
    return NULL;
}
void orc_trace_paths1d(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, int64_t n,
                       int32_t max_bounces, int32_t phase, uint32_t bounce_seed, float box,
                       const float *light, orc_path *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; path_job_t jb[256];
    for (int t = 0; t < nthreads; ++t) {
        jb[t].d = desc; jb[t].rays = rays; jb[t].ls = ls; jb[t].out = out;
        jb[t].max_bounces = max_bounces; jb[t].phase = phase; jb[t].bseed = bounce_seed;
        jb[t].box = box; jb[t].light = light;
        jb[t].n0 = n * t / nthreads; jb[t].n1 = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, path_worker, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ===================================================================== */
/* 3D paths: flights chained through cosine-weighted diffuse bounces     */
/* ===================================================================== */
/* The harness's bounce rule of the device path kernel (the reference has no diffuse
 * bounce on this path; the rule is restated operation for operation so that both give
 * the same bits): sin and cos of 2 pi u from exact quadrant and octant reductions and
 * Taylor polynomials on [0, pi/4] in Horner form; correctly rounded square roots and
 * divisions only. */
static void sincos_turn(float u, float *s, float *c) {
    const float x4 = u * 4.0f;
    const float qf = floorf(x4);
    const int q = (int)qf & 3;
    const float r = x4 - qf;
    const int swap = r > 0.5f;
    const float m = swap ? 1.0f - r : r;
    const float h = m * 1.57079632679489662f;
    const float h2 = h * h;
    float ps = -2.50521083854417188e-8f;
    ps = ps * h2 + 2.75573192239858907e-6f;
    ps = ps * h2 - 1.98412698412698413e-4f;
    ps = ps * h2 + 8.33333333333333333e-3f;
    ps = ps * h2 - 1.66666666666666667e-1f;
    float sn = ps * h2;
    sn = sn * h;
    sn = sn + h;
    float pc = 2.08767569878680990e-9f;
    pc = pc * h2 - 2.75573192239858907e-7f;
    pc = pc * h2 + 2.48015873015873016e-5f;
    pc = pc * h2 - 1.38888888888888889e-3f;
    pc = pc * h2 + 4.16666666666666667e-2f;
    pc = pc * h2 - 0.5f;
    float cs = pc * h2;
    cs = cs + 1.0f;
    const float S = swap ? cs : sn, C = swap ? sn : cs;
    *s = q == 0 ? S : (q == 1 ? C : (q == 2 ? -S : -C));
    *c = q == 0 ? C : (q == 1 ? -S : (q == 2 ? -C : S));
}
void orc_sincos_turn(float u, float *s, float *c) { sincos_turn(u, s, c); }

static v3f unit_normal(v3f g) {
    float l2 = g.x * g.x + g.y * g.y;
    l2 = l2 + g.z * g.z;
    const float inv = 1.0f / sqrtf(l2);
    return V3(g.x * inv, g.y * inv, g.z * inv);
}
static v3f diffuse_direction(uint32_t ray, uint32_t bounce, uint32_t bseed, v3f n, float *jit) {
    pcg_t g;
    pcg_set(&g, (uint64_t)(uint32_t)(orc_xxhash32_4(ray, bounce, bseed, 0u) + 1u));
    const float u1 = pcg_1d(&g), u2 = pcg_1d(&g);
    *jit = pcg_1d(&g);
    const float rr = sqrtf(u1);
    float sn, cs;
    sincos_turn(u2, &sn, &cs);
    const float lx = rr * cs, ly = rr * sn;
    const float lz = sqrtf(fmaxf(0.f, 1.f - u1));
    const float sgn = copysignf(1.0f, n.z);
    const float aa = -1.0f / (sgn + n.z);
    const float bb = (n.x * n.y) * aa;
    const float tx = 1.0f + ((sgn * n.x) * n.x) * aa, ty = sgn * bb, tz = -sgn * n.x;
    const float bx = bb, by = sgn + (n.y * n.y) * aa, bz = -n.y;
    float x = tx * lx + bx * ly;
    x = x + n.x * lz;
    float y = ty * lx + by * ly;
    y = y + n.y * lz;
    float z = tz * lx + bz * ly;
    z = z + n.z * lz;
    float l2 = x * x + y * y;
    l2 = l2 + z * z;
    const float inv = 1.0f / sqrtf(l2);
    return V3(x * inv, y * inv, z * inv);
}
void orc_diffuse_direction(uint32_t ray, uint32_t bounce, uint32_t bseed, const float n[3], float d[3],
                           float *jit) {
    const v3f r = diffuse_direction(ray, bounce, bseed, V3(n[0], n[1], n[2]), jit);
    d[0] = r.x; d[1] = r.y; d[2] = r.z;
}

/* One path: the first flight is the request; every hit continues (up to max_bounces
 * times) from sample.p = origin + t dir (float, each operation rounded) along a diffuse
 * bounce about the unit gradient, to the exit of [-box, box]^3, conditioned on (0, the
 * hit's gradient) with the step-1 sign rule (a secondary flight, prepareGpisSegment
 * SparseConvolutionNoiseMedium.cpp:199-241); an empty segment ends the path as an exit
 * at farT.  `seeded`: flight b uses pixelSampleSegment.w = w0 + b, else the request's. */
static void trace_path3d(const orc_field *f, const orc_ray *ray0, uint32_t path, int max_bounces, uint32_t bseed,
                         float box, int seeded, float perturb, uint32_t pseed, orc_path *out) {
    memset(out, 0, sizeof *out);
    orc_ray ray = *ray0;
    for (int b = 0;; ++b) {
        orc_flight fl;
        trace_flight_impl(f, &ray, &fl, perturb, pseed ^ (path * 0x9E3779B9u) ^ ((uint32_t)b << 24));
        out->flights += 1;
        out->evals += fl.evals;
        if (b == 0) {
            out->t0 = (float)fl.t;
            out->grad0[0] = fl.grad[0]; out->grad0[1] = fl.grad[1]; out->grad0[2] = fl.grad[2];
            out->outcome0 = fl.outcome; out->fail0 = fl.fail_reason;
        }
        out->t_last = (float)fl.t;
        out->outcome_last = fl.outcome; out->fail_last = fl.fail_reason;
        out->pos_last[0] = ray.origin[0]; out->pos_last[1] = ray.origin[1]; out->pos_last[2] = ray.origin[2];
        out->dir_last[0] = ray.dir[0]; out->dir_last[1] = ray.dir[1]; out->dir_last[2] = ray.dir[2];
        if (fl.outcome != ORC_HIT) return;
        out->hits += 1;
        if (b >= max_bounces) return;
        const float t = (float)fl.t;
        const v3f o = V3(ray.origin[0], ray.origin[1], ray.origin[2]), d = V3(ray.dir[0], ray.dir[1], ray.dir[2]);
        const v3f p = V3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
        const v3f g = V3(fl.grad[0], fl.grad[1], fl.grad[2]);
        float jit;
        const v3f w = diffuse_direction(path, (uint32_t)(b + 1), bseed, unit_normal(g), &jit);
        orc_ray nx = *ray0;
        nx.origin[0] = p.x; nx.origin[1] = p.y; nx.origin[2] = p.z;
        nx.dir[0] = w.x; nx.dir[1] = w.y; nx.dir[2] = w.z;
        nx.near_t = 0.f;
        nx.far_t = box_exit_f(p, w, box);
        nx.jitter = jit;
        nx.first_scatter = 0;
        nx.last_val = 0.f;
        nx.last_aniso[0] = g.x; nx.last_aniso[1] = g.y; nx.last_aniso[2] = g.z;
        if (seeded) nx.pss[3] = ray0->pss[3] + (uint32_t)(b + 1);
        if (nx.far_t <= nx.near_t) {
            out->flights += 1;
            out->t_last = nx.far_t;
            out->outcome_last = ORC_EXIT; out->fail_last = ORC_FAIL_NONE;
            return;
        }
        ray = nx;
    }
}
typedef struct {
    const orc_field *f; const orc_ray *rays; orc_path *out; int max_bounces; uint32_t bseed; float box;
    int seeded; float perturb; uint32_t pseed; int64_t n0, n1;
} path3d_job_t;
static void *path3d_worker(void *arg) {
    path3d_job_t *j = (path3d_job_t *)arg;
    for (int64_t i = j->n0; i < j->n1; ++i)
        trace_path3d(j->f, &j->rays[i], (uint32_t)i, j->max_bounces, j->bseed, j->box, j->seeded, j->perturb, j->pseed,
                     &j->out[i]);
    return NULL;
}
void orc_trace_paths3d(const orc_field *f, const orc_ray *rays, int64_t n, int32_t max_bounces, uint32_t bounce_seed,
                       float box, int32_t seeded, float perturb, uint32_t pseed, orc_path *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; path3d_job_t jb[256];
    for (int t = 0; t < nthreads; ++t) {
        jb[t].f = f; jb[t].rays = rays; jb[t].out = out; jb[t].max_bounces = max_bounces; jb[t].bseed = bounce_seed;
        jb[t].box = box; jb[t].seeded = seeded; jb[t].perturb = perturb; jb[t].pseed = pseed;
        jb[t].n0 = n * t / nthreads; jb[t].n1 = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, path3d_worker, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
