// gpis_hip.hip — gfx950 kernels and the C ABI (include/gpis_hip.h) of the
// sparse-convolution GPIS flight engine.
//
// Flight semantics restated from the reference:
//   prepareGpisSegment / finishGpisSegment   src/core/media/SparseConvolutionNoiseMedium.cpp:199-347
//   beginDistance / resumeDistance           src/core/media/GaussianProcessMedium.cpp:236-369
//   conditioning3D                           src/core/math/SparseConvolutionNoise.cpp:498-574
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include <random>
#include <vector>

#include "../../include/gpis_hip.h"
#include "../../include/gpis_surrogate.h"
#include "gpis_1d.hpp"
#include "gpis_field.hpp"
#include "gpis_mlp.hpp"

using namespace gpis;

// ------------------------------------------------------------ contexts --
struct gpis_ctx {
    gpis_field_desc desc;
    FieldParams P;
    FieldParams *P_dev = nullptr;  // device copy of P, read by every kernel (sync_params)
    float *table = nullptr;  // packed impulse table (3N floats per cell)
    size_t table_cells = 0;
    int table_lo[3] = {0, 0, 0}, table_dim[3] = {0, 0, 0};
    // support lists of the table (gpis_lists.hpp): fixed-point impulses, entries, ranges
    int32_t *lfix = nullptr;
    uint32_t *lent = nullptr;
    uint32_t *lrange = nullptr;
    int lists_sub = 3;     // sub-cells per cell edge; 0: no lists
    int lists_points = 2;  // queries one list pass serves (1, or 2 with lists of lists::pair_reach)
    unsigned long long *counters = nullptr;  // device: flights evals steps hits exits fails invalid launches
    unsigned int *work = nullptr;            // device work counter of the persistent kernels
    int device = 0;
    int num_cus = 0;
    int waves_per_cu = 0;
    int block_size = 256;
    int batch_k = 8;  // finish/bounce/refill batch of the 3D kernel (lanes per wave)
    int batch_k_1d = 16;  // ... and of the 1D kernel
    float nf1d = 1.0f;  // sqrt(sparseConvNoiseVariance1D) of the 1D grid
    int fast1d = 1;     // 1D arithmetic: FP32 fast path (1) or the reference's rounding (0)
    float *grid = nullptr;  // device voxels of a tabulated mean
    sdf::MeanParams sdf_host;             // a procedural mean's parameters ...
    sdf::MeanParams *sdf_dev = nullptr;   // ... and their device copy
    std::string err;
};

static void free_lists(gpis_ctx *c);
static gpis_status build_lists(gpis_ctx *c);
static gpis_status sync_params(gpis_ctx *c);

// Media whose conditioned flights (Renewal, Renewal+) evaluate generated cells (per-flight
// realizations, never the impulse table) through noise3d_raw one point at a time: the
// nonstationary fields, single-level and multi-resolution.  Their flight kernels (the COND
// instantiations) make the cells wave-cooperatively (cells_coop).
static int coop_eligible(const FieldParams &P) {
    return (P.multires || P.kind == GPIS_COV_PROC_NS) && P.activate && !P.single && !P.iso3d && !P.full &&
           P.N >= 1 && P.N <= kCoopMaxN;
}

static gpis_status fail(gpis_ctx *c, gpis_status s, const char *msg) {
    if (c) c->err = msg;
    return s;
}
static gpis_status hip_fail(gpis_ctx *c, hipError_t e, const char *where) {
    if (c) c->err = std::string(where) + ": " + hipGetErrorString(e);
    return GPIS_DEVICE_ERROR;
}
#define HIP_TRY(c, expr)                                  \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hip_fail(c, _e, #expr); \
    } while (0)

// The kernels read the field parameters from the context's device copy; every call that
// changes c->P copies it again before returning (synchronous).
static gpis_status sync_params(gpis_ctx *c) {
    if (!c->P_dev) return GPIS_OK;
    HIP_TRY(c, hipDeviceSynchronize());
    HIP_TRY(c, hipMemcpy(c->P_dev, &c->P, sizeof(FieldParams), hipMemcpyHostToDevice));
    return GPIS_OK;
}

// Host restatement of floor(log(s)/log(2.5)) in float (SparseConvolutionNoise.cpp:309).
static int host_level_index(float s) { return (int)floorf(logf(s) / logf(2.5f)); }

// Mat4f::rotAxis((1, 0, 0), deg) rows 0-2 (Mat4f.cpp:132-148), in float with the C
// library's sinf/cosf.
static void rot_axis_x(float deg, float r[9]) {
#pragma clang fp contract(off)
    const float x = 1.0f, y = 0.0f, z = 0.0f;
    const float a = deg * (3.1415926536f / 180.0f);
    const float sn = sinf(a), cs = cosf(a), c1 = 1.0f - cs;
    const float m[9] = {cs + x * x * c1,     x * y * c1 - z * sn, x * z * c1 + y * sn,
                        y * x * c1 + z * sn, cs + y * y * c1,     y * z * c1 - x * sn,
                        z * x * c1 - y * sn, z * y * c1 + x * sn, cs + z * z * c1};
    for (int i = 0; i < 9; ++i) r[i] = m[i];
}

// The procedural SDF mean's parameters (ProceduralMean::fromJson and prepareSdfQueries,
// GPFunctions.cpp:250-266, ProceduralMeanSdf.cpp:11-23), kept on the host until
// gpis_create copies them to the device.
static gpis_status build_sdf_mean(gpis_ctx *c) {
    const gpis_field_desc &d = c->desc;
    sdf::MeanParams &M = c->sdf_host;
    memset(&M, 0, sizeof M);
    if (d.sdf_function < GPIS_SDF_KNOB || d.sdf_function > GPIS_SDF_THREE_SLABS)
        return fail(c, GPIS_INVALID_ARGUMENT, "procedural mean: unknown SDF function");
    M.fn = d.sdf_function;
    M.queries = sdf::supports_queries(M.fn) ? 1 : 0;
    double bound = 0.0;
    for (int i = 0; i < 3; ++i) {
        M.t[i] = (double)d.sdf_inverse_transform[4 * i + 3];
        double row = 0.0;
        for (int j = 0; j < 3; ++j) {
            M.w[3 * i + j] = (double)d.sdf_inverse_transform[4 * i + j];
            row += fabs(M.w[3 * i + j]);
        }
        bound = fmax(bound, row);
    }
    for (int i = 0; i < 12; ++i) M.inv[i] = d.sdf_inverse_transform[i];
    M.minv = d.sdf_min;
    M.scale = d.sdf_scale;
    M.offset = d.sdf_offset;
    // prepareSdfQueries refuses a non-finite or zero transform, scale, offset or minimum
    // for the functions with queries
    if (M.queries) {
        bool finite = std::isfinite(d.sdf_min) && std::isfinite(d.sdf_scale) && std::isfinite(d.sdf_offset);
        for (int i = 0; i < 12; ++i) finite = finite && std::isfinite(d.sdf_inverse_transform[i]);
        if (!finite || !(bound > 0.0))
            return fail(c, GPIS_INVALID_ARGUMENT,
                        "procedural SDF requires a finite invertible transform, scale, offset and minimum");
    }
    rot_axis_x(90.0f, M.rot_p90);
    rot_axis_x(-90.0f, M.rot_m90);
    rot_axis_x(-45.0f, M.rot_m45);
    return GPIS_OK;
}

// 3x3 float inverse by cofactors and one reciprocal of the determinant (the closed form of
// Eigen's size-3 inverse); returns the determinant.
static float inverse3f(const float m[3][3], float r[3][3]) {
    const float c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const float c10 = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    const float c20 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    const float det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    const float inv = 1.0f / det;
    r[0][0] = c00 * inv; r[0][1] = c10 * inv; r[0][2] = c20 * inv;
    r[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    r[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    return det;
}

// The full anisotropy matrix of the SE covariance (SquaredExponentialCovariance::fromJson with
// useAnisoMtx, GPFunctions.cpp:675-681): L = l_conv * anisoMtx, W = L^-1, cinv = W^T W.
// splattingKernelRadius takes the largest component of L's row sums (:714-717),
// sparseConvNoiseVariance3D the factor 1 / sqrt(det cinv) (:757-758), getInvCovMtx cinv
// (:800-806), sparseConvNoiseOneOverSecondDerivative the inverse of -2 cinv / 4 (:738-748).
static gpis_status build_aniso_mtx(gpis_ctx *c, float l_conv) {
    const gpis_field_desc &d = c->desc;
    FieldParams &P = c->P;
    if (d.cov_kind != GPIS_COV_SE)
        return fail(c, GPIS_UNSUPPORTED, "anisoMtx is served for the stationary SE kernel");
    if (d.one_d) return fail(c, GPIS_UNSUPPORTED, "anisoMtx is served for 3D sampling");
    float A[3][3], W[3][3], ci[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = l_conv * d.aniso_mtx[3 * i + j];
    const float detA = inverse3f(A, W);
    bool finite = std::isfinite(detA) && detA != 0.0f;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = W[0][i] * W[0][j];
            s += W[1][i] * W[1][j];
            s += W[2][i] * W[2][j];
            ci[i][j] = s;
            finite = finite && std::isfinite(W[i][j]);
        }
    const float cdet = ci[0][0] * (ci[1][1] * ci[2][2] - ci[1][2] * ci[2][1]) -
                       ci[0][1] * (ci[1][0] * ci[2][2] - ci[1][2] * ci[2][0]) +
                       ci[0][2] * (ci[1][0] * ci[2][1] - ci[1][1] * ci[2][0]);
    if (!finite || !(cdet > 0.0f) || !std::isfinite(cdet))
        return fail(c, GPIS_INVALID_ARGUMENT, "anisoMtx must be a finite invertible matrix");
    float e[3];
    for (int i = 0; i < 3; ++i) e[i] = (A[i][0] + A[i][1]) + A[i][2];
    P.l_max = fmaxf(fmaxf(e[0], e[1]), e[2]);
    if (!(P.l_max > 0.0f)) return fail(c, GPIS_INVALID_ARGUMENT, "anisoMtx gives a non-positive kernel radius");
    P.R = P.kernel_scale * 1.0f * P.l_max;
    P.invR = 1.0f / P.R;
    const float R2 = P.R * P.R, k = -1.4426950408889634f;
    const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) {
        const float m = ci[ii[q]][jj[q]] * 0.5f;
        P.fc[q] = k * (R2 * (q < 3 ? m : 2.0f * m));
        P.fk[q] = ci[ii[q]][jj[q]] * 0.5f * 0.5f;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P.fg[3 * i + j] = -2.f * P.R * (ci[i][j] * 0.5f);
    const double var = (double)(P.Nf / (P.R * P.R * P.R)) * pow(M_PI, 1.5) * (1.0 / sqrt((double)cdet));
    P.nf_inv = (float)(1.0 / sqrt((double)(float)var));
    // inverse of H = -2 cinv / 4 in double from the float entries of H
    double H[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[i][j] = (double)(-2.f * (ci[i][j] * 0.25f));
    const double h00 = H[1][1] * H[2][2] - H[1][2] * H[2][1], h10 = H[0][2] * H[2][1] - H[0][1] * H[2][2],
                 h20 = H[0][1] * H[1][2] - H[0][2] * H[1][1];
    const double hdet = h00 * H[0][0] + h10 * H[1][0] + h20 * H[2][0];
    const double hi[9] = {h00, h10, h20,
                          H[1][2] * H[2][0] - H[1][0] * H[2][2], H[0][0] * H[2][2] - H[0][2] * H[2][0],
                          H[0][2] * H[1][0] - H[0][0] * H[1][2],
                          H[1][0] * H[2][1] - H[1][1] * H[2][0], H[0][1] * H[2][0] - H[0][0] * H[2][1],
                          H[0][0] * H[1][1] - H[0][1] * H[1][0]};
    for (int q = 0; q < 9; ++q) P.fhinv[q] = hi[q] / hdet;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { P.w2l[3 * i + j] = W[i][j]; P.l2w[3 * i + j] = A[i][j]; }
    P.iso = 0;
    P.full = 1;
    return GPIS_OK;
}

static gpis_status build_params(gpis_ctx *c) {
    const gpis_field_desc &d = c->desc;
    FieldParams &P = c->P;
    memset(&P, 0, sizeof P);
    if (!(d.impulse_density >= 1.0f) || d.impulse_density > 64.0f) return fail(c, GPIS_INVALID_ARGUMENT, "impulse_density must be in [1, 64]");
    if (!(d.length_scale > 0.0f) || !std::isfinite(d.length_scale)) return fail(c, GPIS_INVALID_ARGUMENT, "lengthScale must be positive");
    if (!(d.step_size > 0.0f) || d.min_step == 0) return fail(c, GPIS_INVALID_ARGUMENT, "step_size and min_step must be positive");
    if (d.cov_kind != GPIS_COV_SE && d.cov_kind != GPIS_COV_PROC_NS) return fail(c, GPIS_UNSUPPORTED, "covariance kind");
    if (d.one_d && d.cov_kind != GPIS_COV_SE)
        return fail(c, GPIS_UNSUPPORTED, "1D sampling is served for the stationary SE kernel");
    if (d.mean_kind < GPIS_MEAN_SPHERE || d.mean_kind > GPIS_MEAN_PROCEDURAL) return fail(c, GPIS_UNSUPPORTED, "mean kind");
    if (d.mean_kind == GPIS_MEAN_PROCEDURAL) {
        gpis_status s = build_sdf_mean(c);
        if (s != GPIS_OK) return s;
    }
    if (d.isotropic_3d && !d.one_d) {
        if (d.cov_kind != GPIS_COV_SE)
            return fail(c, GPIS_UNSUPPORTED, "isotropic_3D_sampling is served for the stationary SE kernel");
        if (d.mean_kind == GPIS_MEAN_GRID || d.mean_kind == GPIS_MEAN_PROCEDURAL)
            return fail(c, GPIS_UNSUPPORTED, "isotropic_3D_sampling with a tabulated or procedural mean is not served");
        // splattingKernelRadius(true, 1) and sparseConvNoiseVariance3D(N, R, identity)
        // (GPFunctions.cpp:706-719, 751-764); the identity-space splat exp(-|R to|^2 / 2)
        P.iso3d = 1;
        P.iso_R = d.kernel_scale;
        P.iso_c0 = -1.4426950408889634f * (P.iso_R * P.iso_R * 0.5f);
        P.iso_g = -P.iso_R;
        const double var = (double)((float)(unsigned)d.impulse_density / (P.iso_R * P.iso_R * P.iso_R)) * pow(M_PI, 1.5);
        P.iso_nf_inv = (float)(1.0 / sqrt((double)(float)var));
    }
    P.kind = d.cov_kind;
    P.multires = (d.cov_kind == GPIS_COV_PROC_NS) && d.ns_multires;
    P.sigma = d.sigma;
    // SquaredExponentialCovariance::fromJson (GPFunctions.cpp:657-685)
    float l_conv = d.length_scale * sqrtf(2.f) / 2;
    for (int i = 0; i < 3; ++i) {
        P.l_aniso[i] = l_conv * d.aniso[i];
        float inv = 1.0f / P.l_aniso[i];
        P.l_aniso_inv[i] = std::isfinite(inv) ? inv : 0.0f;
        P.aniso[i] = d.aniso[i];
    }
    P.length_scale = d.length_scale;
    P.l_max = fmaxf(fmaxf(P.l_aniso[0], P.l_aniso[1]), P.l_aniso[2]);
    P.kernel_scale = d.kernel_scale;
    P.ls_field = d.ns_ls_field;
    for (int i = 0; i < 4; ++i) P.ls_params[i] = d.ns_ls_params[i];
    P.ls_max = d.cov_kind == GPIS_COV_SE ? 1.0f : d.ns_ls_max;
    P.ns_var = d.ns_var;
    P.amp_const = d.cov_kind == GPIS_COV_SE ? d.sigma : (float)((double)d.ns_var * d.sigma);
    // ProceduralNoise fields: _scale = 1 / (end - start), _offset = -start * _scale
    // (GPFunctions.hpp:726-729)
    auto noise = [](int type, const double *q) {
        NoiseField F;
        F.type = type;
        F.mn = q[0]; F.mx = q[1]; F.scale = 1.0 / (q[3] - q[2]); F.offset = -q[2] * F.scale;
        F.mn2 = q[4]; F.mx2 = q[5]; F.scale2 = 1.0 / (q[7] - q[6]); F.offset2 = -q[6] * F.scale2;
        return F;
    };
    if (d.cov_kind == GPIS_COV_PROC_NS && d.ns_var_field == GPIS_VAR_NOISE) {
        if (d.ns_var_noise < GPIS_NOISE_BOTTOM_TOP || d.ns_var_noise > GPIS_NOISE_RUST)
            return fail(c, GPIS_INVALID_ARGUMENT, "unknown variance noise type");
        P.var_field = 1;
        P.var_noise = noise(d.ns_var_noise, d.ns_var_noise_params);
    } else if (d.ns_var_field != GPIS_VAR_CONST) {
        return fail(c, GPIS_INVALID_ARGUMENT, "variance fields belong to proc_nonstationary covariances");
    }
    if (d.cov_kind == GPIS_COV_PROC_NS && d.ns_ls_field == GPIS_LS_NOISE) {
        if (d.ns_ls_noise < GPIS_NOISE_BOTTOM_TOP || d.ns_ls_noise > GPIS_NOISE_RUST)
            return fail(c, GPIS_INVALID_ARGUMENT, "unknown length-scale noise type");
        P.ls_noise = noise(d.ns_ls_noise, d.ns_ls_noise_params);
    }
    if (d.ns_ls_field < GPIS_LS_CONST || d.ns_ls_field > GPIS_LS_NOISE)
        return fail(c, GPIS_INVALID_ARGUMENT, "unknown length-scale field");
    P.N = (int)(unsigned)d.impulse_density;
    P.iso = d.aniso[0] == d.aniso[1] && d.aniso[1] == d.aniso[2];
    P.Nf = (float)P.N;
    // single-level grid (stationary SE: kss = 1)
    float kss = d.cov_kind == GPIS_COV_SE ? 1.0f : P.ls_max;
    P.seed_add_single = host_level_index(kss);
    for (int L = -32; L <= 32; ++L) P.level_seed_add[L + 32] = host_level_index(powf(2.5f, (float)L));
    P.R = P.kernel_scale * 1.0f * P.l_max;
    P.invR = 1.0f / P.R;
    for (int i = 0; i < 3; ++i) {
        float m = P.l_aniso_inv[i] * P.l_aniso_inv[i] * 0.5f;
        P.c2[i] = -1.4426950408889634f * (P.R * P.R * m);  // exp(-x) = exp2(-log2(e) x)
        P.gfac[i] = -2.f * P.R * m;
    }
    double var = (double)(P.Nf / (P.R * P.R * P.R)) * pow(M_PI, 1.5) *
                 (double)(P.l_aniso[0] * P.l_aniso[1] * P.l_aniso[2]);
    P.nf_inv = (float)(1.0 / sqrt((double)(float)var));
    if (d.use_aniso_mtx) {
        gpis_status s = build_aniso_mtx(c, l_conv);
        if (s != GPIS_OK) return s;
    }
    P.pair_cells = c->lists_points == 2;
    P.mean_kind = d.mean_kind;
    for (int i = 0; i < 3; ++i) {
        P.mc[i] = (float)d.mean_center[i];
        P.mh[i] = (float)d.mean_half[i];
        P.pn[i] = (float)d.plane_normal[i];
    }
    P.mr = d.mean_radius;
    P.poff = (float)d.plane_offset;
    for (int i = 0; i < 3; ++i) {
        P.mcd[i] = d.mean_center[i];
        P.mhd[i] = d.mean_half[i];
        P.pnd[i] = d.plane_normal[i];
    }
    P.poffd = d.plane_offset;
    P.mrd = (double)d.mean_radius;
    P.global_seed = d.global_seed;
    P.single = d.single_realization;
    P.ctxt = d.ctxt;
    P.activate = !d.single_realization && (d.ctxt == GPIS_CTX_RENEWAL || d.ctxt == GPIS_CTX_RENEWAL_PLUS);
    P.one_d = d.one_d;
    P.corr_xy = d.correlation_xy;
    // samplingScheme's configured scheme (SparseConvolutionNoise.cpp:25-54): UNI unless 1D and not single
    P.scheme1d = (!d.single_realization && d.one_d) ? d.scheme : GPIS_SCHEME_UNI;
    P.step_size = d.step_size;
    P.min_step_f = (float)d.min_step;
    P.max_bounce = d.max_bounce;
    P.table = c->table;
    P.has_table = 0;
    // sparseConvNoiseVariance1D(N, R = localScale, 1) (GPFunctions.cpp:766-770)
    c->nf1d = sqrtf((float)((double)(P.Nf / d.kernel_scale) * sqrt(M_PI) * 1.0));
    return GPIS_OK;
}

// ------------------------------------------------------------- kernels --
// Impulses of every cell of a box (bit-exact cell3D RNG). One thread per cell.
__global__ void impulse_table_kernel(int lox, int loy, int loz, int dx, int dy, int dz, uint32_t seed, int n,
                                     float4 *out) {
    long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)dx * dy * dz;
    if (cell >= total) return;
    int x = (int)(cell % dx);
    int y = (int)((cell / dx) % dy);
    int z = (int)(cell / ((long long)dx * dy));
    int cx = lox + x, cy = loy + y, cz = loz + z;
    Pcg g;
    g.set((uint64_t)(uint32_t)(xxh4((uint32_t)cz, (uint32_t)cy, (uint32_t)cx, seed) + 1u));
    float4 *q = out + cell * n;
    for (int k = 0; k < n; ++k) {
        float4 im;
        im.z = u01(g.next());
        im.y = u01(g.next());
        im.x = u01(g.next());
        im.w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
        q[k] = im;
    }
}

// The packed impulse table of a box of cells: per cell N x (x with the weight in its
// sign bit, y, z).  Bit-exact impulse positions; one thread per cell.
__global__ void table_build_kernel(int lox, int loy, int loz, int dx, int dy, int dz, uint32_t seed, int n,
                                   float *out) {
    long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)dx * dy * dz;
    if (cell >= total) return;
    int x = (int)(cell % dx);
    int y = (int)((cell / dx) % dy);
    int z = (int)(cell / ((long long)dx * dy));
    Pcg g;
    g.set((uint64_t)(uint32_t)(xxh4((uint32_t)(loz + z), (uint32_t)(loy + y), (uint32_t)(lox + x), seed) + 1u));
    float *q = out + cell * 3 * n;
    for (int k = 0; k < n; ++k) {
        float iz = u01(g.next());
        float iy = u01(g.next());
        float ix = u01(g.next());
        float w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
        q[3 * k] = copysignf(ix, w);
        q[3 * k + 1] = iy;
        q[3 * k + 2] = iz;
    }
}

// Fixed-point copy of the packed table (gpis_lists.hpp): impulse i at lfix[4i..4i+3], the
// sentinel after the last impulse.  One thread per impulse.
__global__ void fix_build_kernel(const float *__restrict__ table, int lox, int loy, int loz, int dx, int dy,
                                 int n, long long total, int32_t *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > total) return;
    if (i == total) {
        lists::sentinel_impulse(out + 4 * i);
        return;
    }
    const long long cell = i / n;
    const int x = (int)(cell % dx), y = (int)((cell / dx) % dy), z = (int)(cell / ((long long)dx * dy));
    lists::fix_impulse(lox + x, loy + y, loz + z, table[3 * i], table[3 * i + 1], table[3 * i + 2], out + 4 * i);
}

// Support lists, one thread per sub-cell of the table: first the number of groups of four
// entries (0 for cells on the table's border), then, with the groups' positions, the entries.
__global__ void list_count_kernel(const float *__restrict__ table, int lox, int loy, int loz, int dx, int dy, int dz,
                                  int n, int S, double reach, long long nsub, uint32_t *__restrict__ groups) {
    const long long sc = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= nsub) return;
    const int lo[3] = {lox, loy, loz}, dim[3] = {dx, dy, dz};
    const long long cell = sc / (S * S * S);
    const int s = (int)(sc % (S * S * S));
    const int x = lox + (int)(cell % dx), y = loy + (int)((cell / dx) % dy), z = loz + (int)(cell / ((long long)dx * dy));
    uint32_t g = 0;
    if (lists::interior(lo, dim, x, y, z, lists::list_margin(reach)))
        g = (uint32_t)((lists::sub_cell_list(table, lo, dim, n, S, reach, x, y, z, s % S, (s / S) % S, s / (S * S),
                                             nullptr, 0u) + 3) / 4);
    groups[sc] = g;
}
__global__ void list_fill_kernel(const float *__restrict__ table, int lox, int loy, int loz, int dx, int dy, int dz,
                                 int n, int S, double reach, long long nsub, const uint32_t *__restrict__ range,
                                 uint32_t *__restrict__ ent, uint32_t sentinel_off) {
    const long long sc = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= nsub) return;
    if (range[2 * sc + 1] == 0) return;
    const int lo[3] = {lox, loy, loz}, dim[3] = {dx, dy, dz};
    const long long cell = sc / (S * S * S);
    const int s = (int)(sc % (S * S * S));
    const int x = lox + (int)(cell % dx), y = loy + (int)((cell / dx) % dy), z = loz + (int)(cell / ((long long)dx * dy));
    lists::sub_cell_list(table, lo, dim, n, S, reach, x, y, z, s % S, (s / S) % S, s / (S * S),
                         ent + 4 * (size_t)range[2 * sc], sentinel_off);
}

template <int NI, bool ISO, int MEAN>
__global__ __launch_bounds__(256) void eval_points_kernel(const FieldParams *__restrict__ Pg, const float *__restrict__ x,
                                                          const float *__restrict__ y, const float *__restrict__ z,
                                                          long long n, uint32_t seed, float *__restrict__ v,
                                                          float *__restrict__ gx, float *__restrict__ gy,
                                                          float *__restrict__ gz) {
    const FieldParams &P = *Pg;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RealState rs = {};
    rs.seed = seed;
    Eval4 e = eval_field<NI, ISO, MEAN>(P, rs, x[i], y[i], z[i], false);
    v[i] = e.v;
    gx[i] = e.gx;
    gy[i] = e.gy;
    gz[i] = e.gz;
}

// Lane phases of the persistent flight state machine.  Every loop iteration a
// live lane performs exactly one field evaluation at (qx,qy,qz).
enum Phase : int {
    PH_NEED = 0,   // wants a new ray
    PH_COND = 1,   // unconditioned value+gradient at the origin (conditioning)
    PH_F0 = 2,     // f(nearT)
    PH_MARCH = 3,  // lattice step
    PH_SECANT = 4, // secant candidate with 0.9 shrink
    PH_FAR = 5,    // f(farT) on exit
    PH_GRAD = 6,   // gradient at the endpoint
    PH_DEAD = 7,   // no more work
    PH_NEE = 8,    // next-event estimation at the scattering point (1D flights)
    PH_DONE = 9    // flight finished, waiting for the batched finish/bounce/refill step
};

// Waves per SIMD the register budget of the 3D flight kernel must allow.  3 (at most 168
// VGPRs) keeps the 27-cell loop free of spills; measured on the camera-ray paths workload:
// 298.9 M flights/s at 3, 298.2 at 2, 285.4 at 4 (64 bytes of spills), 270.6 at 5.
#ifndef GPIS_FLIGHT_MIN_WAVES
#define GPIS_FLIGHT_MIN_WAVES 3
#endif
// The same for the 1D flight kernels (their FP64 splat exponentials need ~185 VGPRs; at 3
// waves per SIMD they spill and the renewal workload ran 5.72 ms against 5.51).
#ifndef GPIS_FLIGHT1D_MIN_WAVES
#define GPIS_FLIGHT1D_MIN_WAVES 2
#endif

// Lattice steps after which a flight is abandoned and reported invalid.
constexpr int kMaxSteps = 1 << 22;

struct RayIn {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *near_t, *far_t, *jitter;
    const uint32_t *noise_seed, *normal_seed, *flags;
    const float *pv, *pgx, *pgy, *pgz;
};
struct ReplyOut {
    float *t, *value, *gx, *gy, *gz;
    uint32_t *outcome, *evals;
};
struct PathOut {
    float *t0, *gx0, *gy0, *gz0;
    uint32_t *outcome0, *flights, *hits, *evals;
    float *t_last;             // last flight of the path: t and outcome (may be null)
    uint32_t *outcome_last;
};
// Per-flight realization seeds of 3D paths: flight b of path r has pixelSampleSegment
// (px, py, spp, w0 + b) (computeSeed, SparseConvolutionNoise.cpp:57-66); `on` = 0: every
// flight of a path uses the ray's noise_seed.
struct PathSeeds3D {
    const uint32_t *px, *py, *spp, *w0;
    uint32_t scene_hash;  // xxhash32(sceneSeed)
    int on;
};
__device__ __forceinline__ uint32_t flight_seed3d(const FieldParams &P, const PathSeeds3D &ps, unsigned r,
                                                  uint32_t w) {
    const uint32_t a = ps.px[r], b = ps.py[r], s = ps.spp[r];
    uint32_t noise = P.global_seed;
    if (!P.single) noise += (P.ctxt == GPIS_CTX_GLOBAL ? xxh4(a, b, s, 0u) : xxh4(a, b, s, w)) + ps.scene_hash;
    return noise;
}

__device__ __forceinline__ unsigned lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ unsigned mask_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Exact-rounding helpers: each operation is rounded on its own, as in the reference
// build (no fused multiply-add), so lattice points and positions match it bit for bit.
__device__ __forceinline__ float mul_add_exact(float a, float b, float c) {
#pragma clang fp contract(off)
    float m = a * b;
    return c + m;
}

// Position p + t*d evaluated as the reference does (Vec3d arithmetic, then Vec3f):
// one rounded double multiply and add per axis.
__device__ __forceinline__ void ray_point(float ox, float oy, float oz, float dx, float dy, float dz, double t,
                                          float &qx, float &qy, float &qz) {
#pragma clang fp contract(off)
    double mx = t * (double)dx, my = t * (double)dy, mz = t * (double)dz;
    qx = (float)((double)ox + mx);
    qy = (float)((double)oy + my);
    qz = (float)((double)oz + mz);
}
// Gradient position ro + rd*t with rd the double-normalised direction
// (GaussianProcessMedium.cpp:281, :307).
__device__ __forceinline__ void grad_point(float ox, float oy, float oz, float dx, float dy, float dz, double t,
                                           float &qx, float &qy, float &qz) {
#pragma clang fp contract(off)
    double xx = (double)dx * (double)dx, yy = (double)dy * (double)dy, zz = (double)dz * (double)dz;
    double l2 = (xx + yy) + zz;
    double inv = 1.0 / sqrt(l2);
    double nx = (double)dx * inv, ny = (double)dy * inv, nz = (double)dz * inv;
    double mx = nx * t, my = ny * t, mz = nz * t;
    qx = (float)((double)ox + mx);
    qy = (float)((double)oy + my);
    qz = (float)((double)oz + mz);
}
// lerp(a, b, r) = a*(1 - r) + b*r in double (MathUtil.hpp:91-94).
__device__ __forceinline__ double lerp_d(double a, double b, double r) {
#pragma clang fp contract(off)
    double u = a * (1.0 - r);
    double v = b * r;
    return u + v;
}
// Lattice point t0 + k*step in double (exact: both terms fit in 53 bits).
// ray_point, or grad_point when `unit` (the one position computation of a loop
// iteration; the normalisation runs only for the lanes that need it).
__device__ __forceinline__ void point_at(float ox, float oy, float oz, float dx, float dy, float dz, double t, bool unit,
                                         float &qx, float &qy, float &qz) {
#pragma clang fp contract(off)
    double ex = (double)dx, ey = (double)dy, ez = (double)dz;
    if (unit) {
        double xx = ex * ex, yy = ey * ey, zz = ez * ez;
        double l2 = (xx + yy) + zz;
        double inv = 1.0 / sqrt(l2);
        ex = ex * inv; ey = ey * inv; ez = ez * inv;
    }
    double mx = ex * t, my = ey * t, mz = ez * t;
    qx = (float)((double)ox + mx);
    qy = (float)((double)oy + my);
    qz = (float)((double)oz + mz);
}
__device__ __forceinline__ double lattice_t(float t0, int k, float step) {
#pragma clang fp contract(off)
    double m = (double)k * (double)step;
    return (double)t0 + m;
}

// Exit distance of the ray o + t d from the box [-b, b]^3 (slab method).
__device__ __forceinline__ float box_exit(float ox, float oy, float oz, float dx, float dy, float dz, float b) {
    float tx = dx != 0.f ? fmaxf((b - ox) / dx, (-b - ox) / dx) : 3.4e38f;
    float ty = dy != 0.f ? fmaxf((b - oy) / dy, (-b - oy) / dy) : 3.4e38f;
    float tz = dz != 0.f ? fmaxf((b - oz) / dz, (-b - oz) / dz) : 3.4e38f;
    return fmaxf(fminf(tx, fminf(ty, tz)), 0.f);
}

// Conditioning coefficients at a flight origin (conditioning3D, SparseConvolutionNoise.cpp:498-574)
// in FP64.  With one observation point the value and gradient observations decouple
// (the covariance gradient vanishes at zero lag), so the Schur complement of the
// value+gradient observation covariance is the diagonal second-derivative block.
template <int MEAN, bool IRAY, bool NOISE = false>
__device__ __noinline__ void condition_coefficients(const FieldParams &P, const Eval4 &e, float ox, float oy, float oz,
                                                    float pv, float pgx, float pgy, float pgz, float dx, float dy,
                                                    float dz, RealState &rs) {
    float mgx, mgy, mgz;
    double mean = mean_value<MEAN>(P, ox, oy, oz, mgx, mgy, mgz);
    double amp = amplitude<NOISE>(P, ox, oy, oz);
    double cur_v = ((double)e.v - mean) / amp;
    double a = ((double)pv - mean) / amp - cur_v;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    double wsum = 1.0, gsum = 1.0;
    if (P.multires) {
        // kernelScaleLevelRatio at the origin: value /= wlo + whi, gradient /= wlo/slo^2 + whi/shi^2
        // (the sums in float, as the weights and scales are floats)
        const LevelRatio k = level_ratio(ls_field<NOISE>(P, ox, oy, oz));
        wsum = k.wlo + k.whi;
        gsum = k.wlo / (k.slo * k.slo) + k.whi / (k.shi * k.shi);
    }
    if (P.ctxt == GPIS_CTX_RENEWAL_PLUS) {
        double del0 = ((double)pgx - mgx) / amp - ((double)e.gx - mgx) / amp;
        double del1 = ((double)pgy - mgy) / amp - ((double)e.gy - mgy) / amp;
        double del2 = ((double)pgz - mgz) / amp - ((double)e.gz - mgz) / amp;
        if (IRAY) {
            // the isotropic ray space (:539-552): -2 frame.toLocal(localToWorld^T delta)
            const Frame3 F = iso_ray_frame(P, dx, dy, dz);
            double l0 = (double)P.l_aniso[0] * del0, l1 = (double)P.l_aniso[1] * del1,
                   l2 = (double)P.l_aniso[2] * del2;
            if (P.full) {
                const float *L = P.l2w;
                l0 = (double)L[0] * del0 + (double)L[3] * del1 + (double)L[6] * del2;
                l1 = (double)L[1] * del0 + (double)L[4] * del1 + (double)L[7] * del2;
                l2 = (double)L[2] * del0 + (double)L[5] * del1 + (double)L[8] * del2;
            }
            b0 = -2.0 * ((double)F.tx * l0 + (double)F.ty * l1 + (double)F.tz * l2);
            b1 = -2.0 * ((double)F.bx * l0 + (double)F.by * l1 + (double)F.bz * l2);
            b2 = -2.0 * ((double)F.nx * l0 + (double)F.ny * l1 + (double)F.nz * l2);
        } else if (NOISE && P.full) {
            // the full matrix: the inverse of the covariance's second derivative -2 cinv / 4
            const double *H = P.fhinv;
            b0 = H[0] * del0 + H[1] * del1 + H[2] * del2;
            b1 = H[3] * del0 + H[4] * del1 + H[5] * del2;
            b2 = H[6] * del0 + H[7] * del1 + H[8] * del2;
        } else {
            // inverse second derivative of the covariance at zero lag: -2 / l_aniso_inv_i^2
            b0 = -2.0 / ((double)P.l_aniso_inv[0] * P.l_aniso_inv[0]) * del0 / gsum;
            b1 = -2.0 / ((double)P.l_aniso_inv[1] * P.l_aniso_inv[1]) * del1 / gsum;
            b2 = -2.0 / ((double)P.l_aniso_inv[2] * P.l_aniso_inv[2]) * del2 / gsum;
            if (P.kind == GPIS_COV_PROC_NS && !P.multires) {
                // the brute-force kernel splats at l(o): gradient_scale *= sqr(lateral scale)
                // (SparseConvolutionNoise.cpp:563-565), the square formed in float
                const float lat = ls_field<NOISE>(P, ox, oy, oz);
                const double s2 = (double)(lat * lat);
                b0 *= s2; b1 *= s2; b2 *= s2;
            }
        }
    }
    rs.ox = ox; rs.oy = oy; rs.oz = oz;
    rs.a = (float)(a / wsum);
    rs.bx = (float)b0; rs.by = (float)b1; rs.bz = (float)b2;
    rs.cond = 1;
}

// Cosine-weighted diffuse direction about the unit normal n (TangentFrame basis).
__device__ __noinline__ void diffuse_bounce(uint32_t ray, int bounce, uint32_t bounce_seed, float nx, float ny,
                                            float nz, float &dx, float &dy, float &dz, float &jit) {
    diffuse_direction(ray, (uint32_t)bounce, bounce_seed, nx, ny, nz, dx, dy, dz, jit);
}

// Block statistics kept in LDS and flushed once per block.
// (counter 7 is the launch count; ST_SPEC counts evaluations made ahead for a point the
// flight did not ask for)
enum { ST_FLIGHTS = 0, ST_EVALS, ST_STEPS, ST_HITS, ST_EXITS, ST_FAILS, ST_INVALID, ST_SPEC = 8, ST_PASSES, ST_PAIRED,
       ST_N };

// The persistent flight kernel.  Lanes run independent per-flight state machines;
// every iteration each live lane performs one field evaluation.  Lanes whose flight
// (or path) finished take the next ray indices from a global counter: one atomic
// per wave, ranks from the ballot mask.  COND: the medium conditions secondary
// flights (Renewal/Renewal+); PATHS: hits continue as diffuse bounces; GRID: the mean
// is evaluated (tabulated or procedural SDF; the analytic means otherwise, whose variants
// stay free of its registers);
// IRAY: the noise lives in each flight's isotropic ray space (isotropic_3D_sampling).
template <int NI, bool ISO, bool COND, bool PATHS, bool GRID, bool IRAY = false>
__global__ __launch_bounds__(256, GPIS_FLIGHT_MIN_WAVES) void flight_kernel(const FieldParams *__restrict__ Pg, RayIn in, ReplyOut out,
                                                                            PathOut pout, PathSeeds3D pseeds,
                                                                            unsigned int n,
                                                                            unsigned int *work,
                                                                            unsigned long long *stats,
                                                                            int max_bounces, uint32_t bounce_seed,
                                                                            float box, int batch_k) {
    const FieldParams &P = *Pg;
    __shared__ unsigned int bstat[ST_N];
    if (threadIdx.x < ST_N) bstat[threadIdx.x] = 0;
    __syncthreads();
    const unsigned lane = lane_id();
    int phase = PH_NEED;
    unsigned ray = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, near_t = 0, far_t = 0, step = 0, t0 = 0;
    double tq_p = 0.0;  // the lane's next evaluation point: t along the ray (unit_p: unit direction)
    bool unit_p = false;
    // a second evaluation point per pass where one pass can serve both: support lists, or two
    // points in one grid cell of the table or the generated cells of a single-level SE field
    // (the isotropic ray space, the full anisotropy matrix and the nonstationary fields
    // evaluate each point on its own)
    const bool pair_evals = !IRAY && (P.lpair2 > 0.0f || (P.pair_cells && P.kind == GPIS_COV_SE && !(NI < 0 && P.full)));
    const bool coop = COND && !IRAY && P.coop_cells;  // the launch gave each wave its cells_coop LDS
    unsigned lane_passes = 0, lane_paired = 0;  // evaluation passes of this lane, passes with two used
    float pf = 0.f, t_end = 0.f, val_end = 0.f;
    int first = 1, sign0 = 1, k = 0, hit = 0;
    double tk = 0, fac = 0, tp = 0, tt = 0;
    float prev_v = 0, prev_gx = 0, prev_gy = 0, prev_gz = 0;
    RealState rs = {};
    unsigned evals_flight = 0;
    int bounce = 0;
    unsigned path_fh = 0;  // flights (low 16 bits) and hits (high 16 bits) of the current path
    unsigned path_evals = 0;
    uint32_t outcome = 0;     // of a finished flight (PH_DONE)
    float gxv = 0.f, gyv = 0.f, gzv = 0.f;

    for (;;) {
        // ---- finish, bounce and refill, batched.  A lane whose flight is done (PH_DONE)
        // or that wants a ray (PH_NEED) waits until at least batch_k lanes of its wave
        // do, or until no lane is marching; the wave then runs this divergent code once
        // for all of them.  batch_k = 1 serves lanes at once, 64 keeps 64-ray tiles together.
        const unsigned long long waiting = __ballot(phase == PH_DONE || phase == PH_NEED);
        const unsigned long long running = __ballot(phase != PH_DONE && phase != PH_NEED && phase != PH_DEAD);
        if (waiting && ((unsigned)__popcll(waiting) >= (unsigned)batch_k || !running)) {
            if (phase == PH_DONE) {
                const unsigned oc = outcome & 0xFFu;
                atomicAdd(&bstat[ST_FLIGHTS], 1u);
                atomicAdd(&bstat[ST_EVALS], evals_flight);
                atomicAdd(&bstat[ST_STEPS], (unsigned)k);
                atomicAdd(&bstat[oc == GPIS_HIT ? ST_HITS : (oc == GPIS_EXIT ? ST_EXITS : ST_FAILS)], 1u);
                if (!PATHS) {
                    out.t[ray] = t_end; out.value[ray] = val_end;
                    out.gx[ray] = gxv; out.gy[ray] = gyv; out.gz[ray] = gzv;
                    out.outcome[ray] = outcome;
                    if (out.evals) out.evals[ray] = evals_flight;
                    evals_flight = 0;
                    phase = PH_NEED;
                } else {
                    if (bounce == 0) {
                        pout.t0[ray] = t_end; pout.outcome0[ray] = outcome;
                        pout.gx0[ray] = gxv; pout.gy0[ray] = gyv; pout.gz0[ray] = gzv;
                    }
                    path_fh += 1u + (oc == GPIS_HIT ? 0x10000u : 0u);
                    path_evals += evals_flight;
                    evals_flight = 0;
                    phase = PH_NEED;
                    if (oc == GPIS_HIT && bounce < max_bounces) {
                        ++bounce;
                        // sample.p = ray.pos() + t * ray.dir() in float
                        float hx = mul_add_exact(t_end, dx, ox);
                        float hy = mul_add_exact(t_end, dy, oy);
                        float hz = mul_add_exact(t_end, dz, oz);
                        float nx, ny, nz, jit;
                        unit_normal(gxv, gyv, gzv, nx, ny, nz);
                        diffuse_bounce(ray, bounce, bounce_seed, nx, ny, nz, dx, dy, dz, jit);
                        ox = hx; oy = hy; oz = hz;
                        near_t = 0.f;
                        far_t = box_exit(ox, oy, oz, dx, dy, dz, box);
                        prev_v = 0.f; prev_gx = gxv; prev_gy = gyv; prev_gz = gzv;
                        first = 0;
                        rs.cond = 0;
                        if (pseeds.on) rs.seed = flight_seed3d(P, pseeds, ray, pseeds.w0[ray] + (uint32_t)bounce);
                        if (far_t <= near_t) {
                            // an empty segment (the hit on the box's boundary): an exit at farT
                            atomicAdd(&bstat[ST_FLIGHTS], 1u);
                            atomicAdd(&bstat[ST_EXITS], 1u);
                            path_fh += 1u;
                            t_end = far_t;
                            outcome = GPIS_EXIT;
                        } else {
                            step = (far_t - near_t) / P.min_step_f;
                            if (P.step_size < step) step = P.step_size;
                            t0 = mul_add_exact(step, jit, near_t);
                            phase = COND && P.activate ? PH_COND : PH_F0;
                            tq_p = 0.0;  // the origin, or f(nearT) at nearT = 0
                            unit_p = false;
                        }
                    }
                    if (phase == PH_NEED) {
                        pout.flights[ray] = path_fh & 0xFFFFu;
                        pout.hits[ray] = path_fh >> 16;
                        if (pout.evals) pout.evals[ray] = path_evals;
                        if (pout.t_last) {
                            pout.t_last[ray] = t_end;
                            pout.outcome_last[ray] = outcome;
                        }
                    }
                }
            }
            // ---- refill: lanes that want a ray take the next indices (one atomic per
            // wave, ranks from the ballot mask)
            const unsigned long long need = __ballot(phase == PH_NEED);
            if (need) {
                unsigned cnt = (unsigned)__popcll(need);
                unsigned leader = (unsigned)(__ffsll((long long)need) - 1);
                unsigned base = 0;
                if (lane == leader) base = atomicAdd(work, cnt);
                base = __shfl(base, leader);
                if (phase == PH_NEED) {
                    unsigned r = base + mask_rank(need);
                    if (r >= n) {
                        phase = PH_DEAD;
                    } else {
                        ray = r;
                        ox = in.ox[r]; oy = in.oy[r]; oz = in.oz[r];
                        dx = in.dx[r]; dy = in.dy[r]; dz = in.dz[r];
                        near_t = in.near_t[r]; far_t = in.far_t[r];
                        float jit = in.jitter[r];
                        rs.seed = PATHS && pseeds.on ? flight_seed3d(P, pseeds, r, pseeds.w0[r]) : in.noise_seed[r];
                        rs.cond = 0;
                        first = (in.flags[r] & 1u) ? 1 : 0;
                        if (COND && !first) { prev_v = in.pv[r]; prev_gx = in.pgx[r]; prev_gy = in.pgy[r]; prev_gz = in.pgz[r]; }
                        bounce = 0; path_fh = 0; path_evals = 0; evals_flight = 0;
                        // beginDistance (GaussianProcessMedium.cpp:244-265)
                        if (!isfinite(far_t)) far_t = near_t + 2000.f;
                        bool bad = !isfinite(ox) || !isfinite(oy) || !isfinite(oz) || !isfinite(dx) || !isfinite(dy) ||
                                   !isfinite(dz) || !isfinite(near_t) || !(jit >= 0.f && jit < 1.f);
                        if (bad) {
                            if (!PATHS) {
                                out.t[r] = far_t; out.value[r] = 0.f; out.gx[r] = out.gy[r] = out.gz[r] = 0.f;
                                out.outcome[r] = GPIS_INVALID;
                                if (out.evals) out.evals[r] = 0;
                            } else {
                                pout.t0[r] = far_t; pout.outcome0[r] = GPIS_INVALID;
                                pout.gx0[r] = pout.gy0[r] = pout.gz0[r] = 0.f;
                                pout.flights[r] = 0; pout.hits[r] = 0;
                                if (pout.evals) pout.evals[r] = 0;
                                if (pout.t_last) { pout.t_last[r] = far_t; pout.outcome_last[r] = GPIS_INVALID; }
                            }
                            atomicAdd(&bstat[ST_INVALID], 1u);
                            phase = PH_NEED;
                        } else if (far_t <= near_t) {
                            // exited without a flight: no RNG draw, no evaluation
                            if (!PATHS) {
                                out.t[r] = far_t; out.value[r] = first ? 0.f : in.pv[r];
                                out.gx[r] = out.gy[r] = out.gz[r] = 0.f;
                                out.outcome[r] = GPIS_EXIT;
                                if (out.evals) out.evals[r] = 0;
                            } else {
                                pout.t0[r] = far_t; pout.outcome0[r] = GPIS_EXIT;
                                pout.gx0[r] = pout.gy0[r] = pout.gz0[r] = 0.f;
                                pout.flights[r] = 0; pout.hits[r] = 0;
                                if (pout.evals) pout.evals[r] = 0;
                                if (pout.t_last) { pout.t_last[r] = far_t; pout.outcome_last[r] = GPIS_EXIT; }
                            }
                            atomicAdd(&bstat[ST_EXITS], 1u);
                            phase = PH_NEED;
                        } else {
                            // finishGpisSegment step rule (SparseConvolutionNoiseMedium.cpp:289-291)
                            step = (far_t - near_t) / P.min_step_f;
                            if (P.step_size < step) step = P.step_size;
                            t0 = mul_add_exact(step, jit, near_t);
                            if (COND && P.activate && !first) {
                                phase = PH_COND;
                                tq_p = 0.0;  // the origin
                            } else {
                                phase = PH_F0;
                                tq_p = (double)near_t;
                            }
                            unit_p = false;
                        }
                    }
                }
            }
        }
        if (!__ballot(phase != PH_DEAD)) break;
        // a lane with no flight to evaluate skips the evaluation, or, where the wave makes
        // its generated cells together, joins it to make cells for the other lanes
        const bool idle = phase == PH_NEED || phase == PH_DEAD || phase == PH_DONE;
        if (idle && !coop) continue;

        // ---- up to two field evaluations per live lane: the point its flight needs next
        // (tq_p along the ray, along the normalised direction when unit_p; the origin in the
        // conditioning phase) and, when the point after it is already determined by the
        // march (the next lattice point, the next secant candidate, the gradient point at
        // farT), that one too: the two share one pass over the impulses.  The second result
        // is used only when the flight's state after the first asks for exactly that point.
        float pa[3], pb[3] = {0.f, 0.f, 0.f};
        point_at(ox, oy, oz, dx, dy, dz, tq_p, unit_p, pa[0], pa[1], pa[2]);
        double spec_t = 0.0;
        int spec_phase = PH_DEAD;
        bool spec_unit = false;
        if (pair_evals) {
            if (phase == PH_F0 && t0 < far_t) {
                spec_t = (double)t0; spec_phase = PH_MARCH;
            } else if (phase == PH_MARCH) {
                const double tn = lattice_t(t0, k + 1, step);
                if (tn < (double)far_t) { spec_t = tn; spec_phase = PH_MARCH; }
            } else if (phase == PH_SECANT) {
                const double fb = fac * 0.9;
                if (fb > 0.01) { spec_t = lerp_d(tk - (double)step, tk, fb); spec_phase = PH_SECANT; }
            } else if (phase == PH_FAR) {
                spec_t = (double)far_t; spec_phase = PH_GRAD; spec_unit = true;
            }
        }
        const bool two = spec_phase != PH_DEAD && !idle;
        if (two) point_at(ox, oy, oz, dx, dy, dz, spec_t, spec_unit, pb[0], pb[1], pb[2]);
        if (idle) pa[0] = pa[1] = pa[2] = 0.f;
        rs.coop = coop ? (idle ? 2 : 1) : 0;
        Eval4 ea, eb;
        eval_field2<NI, ISO, GRID ? MEAN_EXT : MEAN_ANALYTIC, IRAY>(P, rs, pa, pb, two, !COND || phase != PH_COND, dx,
                                                                     dy, dz, ea, eb);
        if (idle) continue;
        // the next query point: t along the ray (unit: along the normalised direction)
        double tq = 0.0;
        bool unit = false;
        // `second`: the result of the point made ahead, which only a flight in the march, the
        // secant or the gradient phase asks for (the other phases' code drops out)
        auto consume = [&](const Eval4 &e, const bool second) {
            ++evals_flight;
            const float fv = e.v;
            if (COND && !second && phase == PH_COND) {
                condition_coefficients<GRID ? MEAN_EXT : MEAN_ANALYTIC, IRAY, (NI < 0)>(P, e, ox, oy, oz, prev_v, prev_gx, prev_gy,
                                                                              prev_gz, dx, dy, dz, rs);
                phase = PH_F0;
                tq = (double)near_t;
            } else if (!second && phase == PH_F0) {
                sign0 = fv < 0.f ? -1 : 1;
                pf = fv;
                k = 0;
                if (t0 < far_t) {
                    phase = PH_MARCH;
                    tq = (double)t0;
                } else {
                    phase = PH_FAR;
                    tq = (double)far_t;
                }
            } else if (phase == PH_MARCH) {
                ++k;
                int sc = fv < 0.f ? -1 : 1;
                // lattice t_k = t0 + k*step, exact in double as the reference's `t += step_size`
                tk = lattice_t(t0, k - 1, step);
                if (!first && k == 1) {
                    sign0 = sc;
                } else if (sc != sign0) {
                    // secant with 0.9 shrink (SparseConvolutionNoiseMedium.cpp:309-329), in FP64
                    fac = (double)pf / ((double)pf - (double)fv);
                    tp = lerp_d(tk - (double)step, tk, fac);
                    tt = tp;
                    phase = PH_SECANT;
                    tq = tt;
                }
                if (phase == PH_MARCH) {
                    pf = fv;
                    double tn = lattice_t(t0, k, step);
                    if (k >= kMaxSteps) {
                        // step limit (RaySegmentStatus::StepLimit, RaySegmentBatch.cu:281-284): reported invalid
                        atomicAdd(&bstat[ST_INVALID], 1u);
                        atomicAdd(&bstat[ST_STEPS], (unsigned)k);
                        if (!PATHS) {
                            out.t[ray] = (float)tn; out.value[ray] = fv;
                            out.gx[ray] = out.gy[ray] = out.gz[ray] = 0.f;
                            out.outcome[ray] = GPIS_INVALID;
                            if (out.evals) out.evals[ray] = evals_flight;
                        } else {
                            if (bounce == 0) { pout.t0[ray] = (float)tn; pout.outcome0[ray] = GPIS_INVALID; }
                        if (pout.t_last) { pout.t_last[ray] = (float)tn; pout.outcome_last[ray] = GPIS_INVALID; }
                            pout.flights[ray] = path_fh & 0xFFFFu; pout.hits[ray] = path_fh >> 16;
                            if (pout.evals) pout.evals[ray] = path_evals + evals_flight;
                        }
                        evals_flight = 0;
                        phase = PH_NEED;
                    } else if (tn < (double)far_t) {
                        tq = tn;
                    } else {
                        phase = PH_FAR;
                        tq = (double)far_t;
                    }
                }
            } else if (phase == PH_SECANT) {
                bool done = false;
                if ((fv < 0.f ? -1 : 1) == sign0) {
                    done = true;
                } else {
                    fac *= 0.9;
                    if (fac <= 0.01) {
                        tp = 0.0;
                        done = true;
                    } else {
                        tp = tt;
                        tt = lerp_d(tk - (double)step, tk, fac);
                        tq = tt;
                    }
                }
                if (done) {
                    hit = 1;
                    t_end = (float)tp;
                    val_end = 0.f;
                    phase = PH_GRAD;
                    tq = tp;
                    unit = true;
                }
            } else if (!second && phase == PH_FAR) {
                hit = 0;
                t_end = far_t;
                val_end = fv;
                phase = PH_GRAD;
                tq = (double)far_t;
                unit = true;
            } else if (phase == PH_GRAD) {
                // resumeDistance outcome rules (GaussianProcessMedium.cpp:305-348); the rest of
                // the finish waits for the batched step at the top of the loop
                gxv = e.gx; gyv = e.gy; gzv = e.gz;
                if (hit) {
                    if (!isfinite(gxv + gyv + gzv)) outcome = GPIS_FAIL | (GPIS_FAIL_NONFINITE << 8);
                    else if (gxv * dx + gyv * dy + gzv * dz > 0.f) outcome = GPIS_FAIL | (GPIS_FAIL_OUTWARD << 8);
                    else if (gxv * gxv + gyv * gyv + gzv * gzv < 1e-7f) outcome = GPIS_FAIL | (GPIS_FAIL_ZERO_GRAD << 8);
                    else outcome = GPIS_HIT;
                } else {
                    outcome = GPIS_EXIT;
                }
                phase = PH_DONE;
            }
        };
        consume(ea, false);
        ++lane_passes;
        if (two) {
            if (phase == spec_phase && tq == spec_t && unit == spec_unit) {
                consume(eb, true);
                ++lane_paired;
            } else {
                atomicAdd(&bstat[ST_SPEC], 1u);
            }
        }
        tq_p = tq;
        unit_p = unit;
    }
    // ---- statistics: LDS block totals, one global atomic per counter per block
    atomicAdd(&bstat[ST_PASSES], lane_passes);
    atomicAdd(&bstat[ST_PAIRED], lane_paired);
    __syncthreads();
    if (threadIdx.x < ST_N && bstat[threadIdx.x]) atomicAdd(&stats[threadIdx.x], (unsigned long long)bstat[threadIdx.x]);
}


// ---------------------------------------------------------------- 1D flights --
struct Nee1DOut {
    float *gx, *gy, *gz, *pdf;
};

// Per-lane 1D field evaluation at q: normalised noise value and derivative along the
// ray (conditioned when `cond`), then f = (float)(amplitude * noise + mean in double),
// the reference's evaluateValue rounding.  FAST: FP32 splats, conditioning splats and
// analytic mean (the normalisation as a multiplication by 1/nf).
template <int NI, bool PROC, bool FAST = false>
__device__ __forceinline__ void eval1d(const FieldParams &P, const Ray1D &R, Cells1D<NI> &C, bool &cache_valid,
                                       float nf, float qx, float qy, float qz, bool cond, float &f, float &nv,
                                       float &ndv) {
#pragma clang fp contract(off)
    float u = ray_coord(R, qx, qy, qz);
    float v, dv;
    noise1d_raw<NI, FAST>(C, cache_valid, u, R.seed, v, dv);
    if (FAST) {
        const float inf = 1.0f / nf;
        v *= inf;
        dv *= inf;
    } else {
        v = v / nf;
        dv = dv / nf;
    }
    if (cond && R.cond) cond_splat1d<FAST>(R, u, v, dv);
    nv = v;
    ndv = dv;
    if (FAST) f = P.amp_const * v + mean_value_f<PROC>(P, qx, qy, qz);
    else f = (float)((double)(P.amp_const * v) + mean_value_d<PROC>(P, qx, qy, qz));
}

// eval1d with the cells in LDS, generated by the wave (flight1d_kernel).
template <int NI, bool PROC, bool FAST = false>
__device__ __forceinline__ void eval1d_shared(const FieldParams &P, const Ray1D &R, Cells1DShared<NI> &C, float nf,
                                              float qx, float qy, float qz, bool cond, float &f, float &nv,
                                              float &ndv) {
#pragma clang fp contract(off)
    float u = ray_coord(R, qx, qy, qz);
    float v, dv;
    noise1d_shared<NI, FAST>(C, u, R.seed, v, dv);
    if (FAST) {
        const float inf = 1.0f / nf;
        v *= inf;
        dv *= inf;
    } else {
        v = v / nf;
        dv = dv / nf;
    }
    if (cond && R.cond) cond_splat1d<FAST>(R, u, v, dv);
    nv = v;
    ndv = dv;
    if (FAST) f = P.amp_const * v + mean_value_f<PROC>(P, qx, qy, qz);
    else f = (float)((double)(P.amp_const * v) + mean_value_d<PROC>(P, qx, qy, qz));
}

// neeShared + neeGrad + neePDF (SparseConvolutionNoise.cpp:662-821) for the single-level
// stationary kernel: the mirror normal `h`, the hit point p, the conditioned derivative
// along the ray `dvp`; with `tr` the transverse density is the correlated law of
// seTransverseGradientMoments (tmx, tmy, tvar).  Float arithmetic as the reference.
template <bool PROC>
__device__ __forceinline__ void nee1d(const FieldParams &P, const Ray1D &R, float hx, float hy, float hz, float px,
                                   float py, float pz, float dvp, float dx, float dy, float dz, bool tr, double tmx,
                                   double tmy, double tvar, float &gx, float &gy, float &gz, float &pdf) {
#pragma clang fp contract(off)
    float mgx, mgy, mgz;
    mean_grad_d<PROC>(P, px, py, pz, mgx, mgy, mgz);
    const float amp = P.amp_const;
    const float s0 = amp * R.ia[0], s1 = amp * R.ia[1], s2 = amp * R.ia[2];
    // M = amp * diag(1/l_aniso) * [tangent bitangent normal]
    float M[3][3] = {{s0 * R.F.tx, s0 * R.F.bx, s0 * R.F.nx},
                     {s1 * R.F.ty, s1 * R.F.by, s1 * R.F.ny},
                     {s2 * R.F.tz, s2 * R.F.bz, s2 * R.F.nz}};
    // Eigen's 3x3 inverse: cofactors scaled by 1/det
    float c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    float c10 = M[0][2] * M[2][1] - M[0][1] * M[2][2];
    float c20 = M[0][1] * M[1][2] - M[0][2] * M[1][1];
    float det = c00 * M[0][0] + c10 * M[1][0] + c20 * M[2][0];
    float idet = 1.0f / det;
    float I[3][3];
    I[0][0] = c00 * idet; I[0][1] = c10 * idet; I[0][2] = c20 * idet;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * idet;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * idet;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * idet;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * idet;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * idet;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * idet;
    // the gradient plane: a point and the two transverse axes
    float ppx = M[0][2] * dvp + mgx, ppy = M[1][2] * dvp + mgy, ppz = M[2][2] * dvp + mgz;
    float ax = M[0][0], ay = M[1][0], az = M[2][0];
    float bx = M[0][1], by = M[1][1], bz = M[2][1];
    float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    float cl2 = cx * cx;
    cl2 += cy * cy;
    cl2 += cz * cz;
    float stretch = sqrtf(cl2);
    float il = 1.0f / stretch;
    float nx = cx * il, ny = cy * il, nz = cz * il;
    float isect = dot3(ppx, ppy, ppz, nx, ny, nz) / dot3(hx, hy, hz, nx, ny, nz);
    gx = isect * hx; gy = isect * hy; gz = isect * hz;
    pdf = 0.0f;
    if (isect < 0.0f) return;
    float rx = gx - mgx, ry = gy - mgy, rz = gz - mgz;
    float g3x = I[0][0] * rx + I[0][1] * ry;
    g3x += I[0][2] * rz;
    float g3y = I[1][0] * rx + I[1][1] * ry;
    g3y += I[1][2] * rz;
    float pdf_area;
    if (tr) {
        // the correlated transverse law (neePDF :742-752): a delta at the origin
        if (!(tvar > 0.0)) { pdf = 0.0f; return; }
        const double ex = (double)g3x - tmx, ey = (double)g3y - tmy;
        pdf_area = (float)(exp(-0.5 * (ex * ex + ey * ey) / tvar) / (2.0 * M_PI * tvar));
    } else {
        // transverse draw density with variance 1/2 per axis (no correlation with the origin)
        float smx = (float)((double)g3x * 1.4142135623730951 * 1.0);
        float smy = (float)((double)g3y * 1.4142135623730951 * 1.0);
        pdf_area = (float)(exp((double)(-smx * smx) / 2.0) * exp((double)(-smy * smy) / 2.0) / (2.0 * M_PI) * 2.0 *
                           1.0 * 1.0);
    }
    float cosl = fabsf(dot3(hx, hy, hz, nx, ny, nz));
    float pdf_n = pdf_area * (isect * isect) / cosl / stretch;
    pdf = (float)(pdf_n / (4.0 * (double)dot3(hx, hy, hz, -dx, -dy, -dz)));
}

// Inputs and outputs of device-resident 1D paths (gpis_trace_paths_1d).
struct Path1DIn {
    const uint32_t *px, *py, *spp, *w0;  // RayInfo::pixelSampleSegment of each path's first flight
    uint32_t scene_hash;                 // xxhash32(sceneSeed)
    int max_bounces, phase;
    uint32_t bounce_seed;
    float box;
};
struct Path1DOut {
    float *t0, *gx0, *gy0, *gz0;
    uint32_t *outcome0, *flights, *hits, *shadows, *visible;
    float *nee, *t_last;
    uint32_t *outcome_last, *evals;
};

// Seeds of the flight with pixelSampleSegment.w = w of path r: computeSeed
// (SparseConvolutionNoise.cpp:57-66) and the normal seed of prepareGpisSegment
// (SparseConvolutionNoiseMedium.cpp:235-238).
__device__ __forceinline__ void path_seeds(const FieldParams &P, const Path1DIn &pin, unsigned r, uint32_t w,
                                           uint32_t &noise, uint32_t &normal) {
    const uint32_t a = pin.px[r], b = pin.py[r], s = pin.spp[r];
    const uint32_t h = xxh4(a, b, s, w);
    normal = h + pin.scene_hash + 1u;
    noise = P.global_seed;
    if (!P.single) noise += (P.ctxt == GPIS_CTX_GLOBAL ? xxh4(a, b, s, 0u) : h) + pin.scene_hash;
}

// The scattering frame of BRDFPhaseFunction::setEventIsectInfo (BRDFPhaseFunction.cpp:29-49):
// TangentFrame(Vec3f(aniso.normalized())), the gradient normalised in double, and
// wi = frame.toLocal(-dir).normalized().
__device__ __forceinline__ void hit_frame(float gx, float gy, float gz, float dx, float dy, float dz, Frame3 &F,
                                          float &wix, float &wiy, float &wiz) {
#pragma clang fp contract(off)
    const double ex = gx, ey = gy, ez = gz;
    double l2 = ex * ex;
    l2 += ey * ey;
    l2 += ez * ez;
    const double inv = 1.0 / sqrt(l2);
    F = make_frame((float)(ex * inv), (float)(ey * inv), (float)(ez * inv));
    wix = dot3(F.tx, F.ty, F.tz, -dx, -dy, -dz);
    wiy = dot3(F.bx, F.by, F.bz, -dx, -dy, -dz);
    wiz = dot3(F.nx, F.ny, F.nz, -dx, -dy, -dz);
    normalize3(wix, wiy, wiz);
}
// TangentFrame::toGlobal: tangent * x + bitangent * y + normal * z, each product rounded.
__device__ __forceinline__ void frame_to_global(const Frame3 &F, float x, float y, float z, float &ox, float &oy,
                                                float &oz) {
#pragma clang fp contract(off)
    const float ax = F.tx * x, ay = F.ty * x, az = F.tz * x;
    const float bx = F.bx * y, by = F.by * y, bz = F.bz * y;
    const float cx = F.nx * z, cy = F.ny * z, cz = F.nz * z;
    ox = (ax + bx) + cx;
    oy = (ay + by) + cy;
    oz = (az + bz) + cz;
}
// Angle.hpp:8-14 in float: PI = 3.1415926536f, TWO_PI = PI * 2.0f, INV_PI = 1.0f / PI
#define kTwoPiF 0x1.921fb6p+2f
#define kInvPiF 0x1.45f306p-2f

// conditioning1D (SparseConvolutionNoise.cpp:576-656) from the unconditioned evaluation at
// the origin (normalised value nv, derivative along the ray ndv): the value coefficient v1
// and, for Renewal+, the derivative coefficient g1z and the observed transverse gradient
// of the correlated law.
template <bool PROC>
__device__ __forceinline__ void condition1d(const FieldParams &P, Ray1D &R, float ox, float oy, float oz, float nv,
                                            float ndv, float prev_v, float prev_gx, float prev_gy, float prev_gz,
                                            float bx1, float by1, float &obs_x, float &obs_y, bool &observed) {
#pragma clang fp contract(off)
    const float amp = P.amp_const;
    R.uo = ray_coord(R, ox, oy, oz);
    R.v1 = 0.f;
    R.g1z = 0.f;
    if (amp != 0.f) {
        const float mean = (float)mean_value_d<PROC>(P, ox, oy, oz);
        R.v1 = (prev_v - mean) / amp - nv;
        if (P.ctxt == GPIS_CTX_RENEWAL_PLUS) {
            float mgx, mgy, mgz;
            mean_grad_d<PROC>(P, ox, oy, oz, mgx, mgy, mgz);
            float cgx, cgy, cgz;
            ray_grad_to_world(R, bx1, by1, ndv, cgx, cgy, cgz);
            const float tgx = (prev_gx - mgx) / amp, tgy = (prev_gy - mgy) / amp, tgz = (prev_gz - mgz) / amp;
            float ddx = tgx - cgx;
            float ddy = tgy - cgy;
            float ddz = tgz - cgz;
            // gradient_scale = -2 * frame.toLocal(l_aniso * delta) * scale^2; only its
            // z component enters the value and the derivative along the ray (the
            // transverse components are the correlated law's, through the moments)
            float gsz = dot3(R.F.nx, R.F.ny, R.F.nz, R.la[0] * ddx, R.la[1] * ddy, R.la[2] * ddz);
            R.g1z = -2.f * gsz * 1.0f;
            // target_gradient_noise in isotropic ray space (seTransverseGradientMoments :94-97)
            const float gwx = R.la[0] * tgx, gwy = R.la[1] * tgy, gwz = R.la[2] * tgz;
            obs_x = dot3(R.F.tx, R.F.ty, R.F.tz, gwx, gwy, gwz);
            obs_y = dot3(R.F.bx, R.F.by, R.F.bz, gwx, gwy, gwz);
            observed = true;
        }
    }
    R.cond = 1;
}

// evaluateGradientNoise1D at the endpoint q (:204-242) — transverse pair 0 when tSegment == 0,
// otherwise an independent second pair, or the correlated law's pair 0 about the conditional
// mean (:186-196) — then resumeDistance's rejection rules in double
// (GaussianProcessMedium.cpp:289-369).  Returns the outcome.
template <bool PROC>
__device__ __forceinline__ uint32_t finish1d(const FieldParams &P, const Ray1D &R, bool corr, bool observed,
                                             float obs_x, float obs_y, float ox, float oy, float oz, float qx,
                                             float qy, float qz, float dx, float dy, float dz, uint32_t normal_seed,
                                             float t_end, float ndv_end, int hit, float &gxv, float &gyv,
                                             float &gzv) {
    float tx2, ty2;
    if (corr) {
        double mx, my, var, z0, z1;
        transverse_moments(R, observed, obs_x, obs_y, ox, oy, oz, qx, qy, qz, mx, my, var);
        normal_pair(normal_seed, 0, z0, z1);
        const double sd = sqrt(var);
        tx2 = (float)(mx + sd * z0);
        ty2 = (float)(my + sd * z1);
    } else {
        double x2, y2;
        transverse_pair(normal_seed, t_end == 0.0f ? 0 : 1, x2, y2);
        tx2 = (float)x2;
        ty2 = (float)y2;
    }
    float wx, wy, wz;
    ray_grad_to_world(R, tx2, ty2, ndv_end, wx, wy, wz);
    float mgx, mgy, mgz;
    mean_grad_d<PROC>(P, qx, qy, qz, mgx, mgy, mgz);
    const float amp = P.amp_const;
    {
#pragma clang fp contract(off)
        gxv = amp * wx + mgx; gyv = amp * wy + mgy; gzv = amp * wz + mgz;
    }
    if (!hit) return GPIS_EXIT;
    const double gx = gxv, gy = gyv, gz = gzv;
    double gd = gx * (double)dx;
    gd += gy * (double)dy;
    gd += gz * (double)dz;
    double g2 = 0.0;
    g2 += gx * gx;
    g2 += gy * gy;
    g2 += gz * gz;
    if (!isfinite((gx + gy + gz) * (1.0 / 3.0))) return GPIS_FAIL | (GPIS_FAIL_NONFINITE << 8);
    if (gd > 0.0) return GPIS_FAIL | (GPIS_FAIL_OUTWARD << 8);
    if (g2 < (double)1e-7f) return GPIS_FAIL | (GPIS_FAIL_ZERO_GRAD << 8);
    return GPIS_HIT;
}

// The 1D persistent flight kernel: same lane state machine as flight_kernel, with the
// 1D field (cells cached per lane), conditioning1D at the origin, the Box-Muller
// transverse gradient at the endpoint, and optional NEE toward a point light.
// PATHS: each lane traces a whole path instead (gpis_trace_paths_1d): at a hit the
// hit-time phase runs in the batched finish step (the light sample with its shadow
// flight, then the mirror or Lambert bounce) and the next flight starts from the hit,
// conditioned on (0, gradient), without leaving the kernel.
template <int NI, bool PATHS, bool PROC, bool FAST>
__global__ __launch_bounds__(256, GPIS_FLIGHT1D_MIN_WAVES) void flight1d_kernel(const FieldParams *__restrict__ Pg, RayIn in, const float *__restrict__ lscale,
                                                          ReplyOut out, unsigned int n, unsigned int *work,
                                                          unsigned long long *stats, float nf, int use_light,
                                                          float lx, float ly, float lz, Nee1DOut nee,
                                                          const uint32_t *__restrict__ idx,
                                                          const uint32_t *__restrict__ n_dev, int batch_k,
                                                          Path1DIn pin, Path1DOut pout) {
    const FieldParams &P = *Pg;
    // idx / n_dev: optional list of ray indices (and its device-side length) to trace
    const unsigned nn = n_dev ? *n_dev : n;
    __shared__ unsigned int bstat[ST_N];
    if (threadIdx.x < ST_N) bstat[threadIdx.x] = 0;
    __syncthreads();
    const unsigned lane = lane_id();
    int phase = PH_NEED;
    unsigned ray = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, near_t = 0, far_t = 0, step = 0, t0 = 0;
    float qx = 0, qy = 0, qz = 0, pf = 0.f, t_end = 0.f, val_end = 0.f;
    int first = 1, sign0 = 1, k = 0, hit = 0;
    double tk = 0, fac = 0, tp = 0, tt = 0;
    float prev_v = 0, prev_gx = 0, prev_gy = 0, prev_gz = 0;
    uint32_t normal_seed = 0;
    unsigned evals_flight = 0;
    float gxv = 0, gyv = 0, gzv = 0;
    uint32_t outcome = 0;
    float bx1 = 0.f, by1 = 0.f;      // transverse gradient of the unconditioned field at the origin
    float obs_x = 0.f, obs_y = 0.f;  // observed transverse gradient at the origin (correlated law)
    bool observed = false;           // coeff_1D.has_gradient_observation
    const bool corr = P.corr_xy && P.ctxt == GPIS_CTX_RENEWAL_PLUS;  // seTransverseGradientMoments applies
    float ndv_end = 0.f;             // derivative along the ray at the endpoint
    float ndv_nee = 0.f;             // ... and at the NEE point
    // 0: flight running or its gradient pending; 1: NEE evaluation done; 2: reply ready
    // (flights); 3: the bounce is next (paths)
    int nee_stage = 0;
    Ray1D R = {};
    // the lane's three cells in LDS (a ring per lane), made by the whole wave
    __shared__ float4 cells_lds[4 * 3 * (NI / 4) * 64];
    __shared__ uint32_t jobs_lds[4 * 2 * Cells1DShared<NI>::kMaxJobs];
    Cells1DShared<NI> C;
    {
        const unsigned w = threadIdx.x / 64u;
        C.cells = cells_lds + w * (3 * (NI / 4) * 64);
        C.job_st = jobs_lds + w * (2 * Cells1DShared<NI>::kMaxJobs);
        C.job_dst = C.job_st + Cells1DShared<NI>::kMaxJobs;
        C.base = 0;
        C.valid = false;
    }
    // path state (PATHS)
    int pb = 0;                 // flight index within the path (pixelSampleSegment.w - w0)
    bool shadow = false;        // the running flight is a light sample's shadow flight
    uint32_t w0 = 0;
    float hx = 0.f, hy = 0.f, hz = 0.f;     // scattering point of the last hit (sample.p)
    float hgx = 0.f, hgy = 0.f, hgz = 0.f;  // its gradient (sample.aniso)
    float mdx = 0.f, mdy = 0.f, mdz = 0.f;  // direction of the flight that hit
    float hnx = 0.f, hny = 0.f, hnz = 0.f;  // the mirror's half-vector normal toward the light
    float nee_f = 0.f, nee_rsq = 1.f, nee_acc = 0.f, t_last = 0.f;
    uint32_t o_last = 0;
    unsigned pfl = 0, phit = 0, psh = 0, pvis = 0, pev = 0;

    // Set up the lane's flight from ox..dz, near_t, far_t (finite, > near_t), first, the
    // previous value/gradient, R.seed and normal_seed: ray frame, step rule and the first
    // phase (conditioning at the origin, or f(nearT)).
    auto begin_flight = [&](float jit) __attribute__((always_inline)) {
        evals_flight = 0;
        nee_stage = 0;
        observed = false;
        C.valid = false;
        R.cond = 0;
        float ix = R.ia[0] * dx, iy = R.ia[1] * dy, iz = R.ia[2] * dz;
        normalize3(ix, iy, iz);
        R.F = make_frame(ix, iy, iz);
        // finishGpisSegment step rule (SparseConvolutionNoiseMedium.cpp:289-291)
        step = (far_t - near_t) / P.min_step_f;
        if (P.step_size < step) step = P.step_size;
        t0 = mul_add_exact(step, jit, near_t);
        if (P.activate && !first) {
            phase = PH_COND;
            qx = ox; qy = oy; qz = oz;
            if (corr) {
                // evaluateGradientNoise1D(conditioning = false): mean 0, variance 1/2
                double z0, z1;
                normal_pair(normal_seed, 0, z0, z1);
                const double sd = sqrt(0.5);
                bx1 = (float)(0.0 + sd * z0);
                by1 = (float)(0.0 + sd * z1);
            } else if (P.ctxt == GPIS_CTX_RENEWAL_PLUS) {
                double x1, y1;
                transverse_pair(normal_seed, 0, x1, y1);
                bx1 = (float)x1; by1 = (float)y1;
            }
        } else {
            phase = PH_F0;
            ray_point(ox, oy, oz, dx, dy, dz, (double)near_t, qx, qy, qz);
        }
    };
    auto end_path = [&]() __attribute__((always_inline)) {
        pout.flights[ray] = pfl; pout.hits[ray] = phit;
        pout.shadows[ray] = psh; pout.visible[ray] = pvis;
        pout.nee[ray] = nee_acc; pout.t_last[ray] = t_last; pout.outcome_last[ray] = o_last;
        if (pout.evals) pout.evals[ray] = pev;
        phase = PH_NEED;
    };
    // A flight that ends before its march (empty interval or medium bounce limit).
    auto no_march_flight = [&](uint32_t oc) __attribute__((always_inline)) {
        if (pb == 0) {
            pout.t0[ray] = far_t; pout.gx0[ray] = pout.gy0[ray] = pout.gz0[ray] = 0.f;
            pout.outcome0[ray] = oc;
        }
        ++pfl;
        t_last = far_t;
        o_last = oc;
        atomicAdd(&bstat[(oc & 0xFFu) == GPIS_EXIT ? ST_EXITS : ST_FAILS], 1u);
        end_path();
    };
    // volumeLightSample's shadow flight (TraceBase.cpp:812-828): from the scattering point
    // toward the light (Point::sampleDirect, Point.cpp:98-106) over min(distance, box
    // exit), conditioned on (0, sg), with this flight's seeds.
    auto start_shadow = [&](float sgx, float sgy, float sgz) __attribute__((always_inline)) {
        float dlx = lx - hx, dly = ly - hy, dlz = lz - hz;
        const float dist = sqrtf(dot3(dlx, dly, dlz, dlx, dly, dlz));
        dlx = dlx / dist; dly = dly / dist; dlz = dlz / dist;
        ox = hx; oy = hy; oz = hz;
        dx = dlx; dy = dly; dz = dlz;
        near_t = 0.f;
        far_t = fminf(dist, box_exit(hx, hy, hz, dlx, dly, dlz, pin.box));
        first = 0;
        prev_v = 0.f; prev_gx = sgx; prev_gy = sgy; prev_gz = sgz;
        Pcg g;
        g.set((uint64_t)(uint32_t)(xxh4(ray, (uint32_t)pb, pin.bounce_seed, 1u) + 1u));
        const float jit = u01(g.next());
        if (far_t <= near_t) {
            // transmittance over an empty segment is 1
            ++psh; ++pvis;
            nee_acc += nee_f / nee_rsq;
            nee_stage = 3;
        } else {
            shadow = true;
            begin_flight(jit);
        }
    };

    for (;;) {
        // ---- finish and refill, batched as in flight_kernel: lanes whose flight ended
        // (PH_DONE) or that want a ray wait until batch_k of them can be served together.
        const unsigned long long waiting = __ballot(phase == PH_DONE || phase == PH_NEED);
        const unsigned long long running = __ballot(phase != PH_DONE && phase != PH_NEED && phase != PH_DEAD);
        if (waiting && ((unsigned)__popcll(waiting) >= (unsigned)batch_k || !running)) {
            if (PATHS && phase == PH_DONE && nee_stage == 0 && shadow) {
                // the shadow flight ended: any crossing occludes (resumeTransmittance,
                // GaussianProcessMedium.cpp:394-421), otherwise f * e / pdf with e = 1
                atomicAdd(&bstat[ST_FLIGHTS], 1u);
                atomicAdd(&bstat[ST_EVALS], evals_flight);
                atomicAdd(&bstat[ST_STEPS], (unsigned)k);
                atomicAdd(&bstat[hit ? ST_HITS : ST_EXITS], 1u);
                pev += evals_flight;
                ++psh;
                if (!hit) {
                    ++pvis;
                    nee_acc += nee_f / nee_rsq;
                }
                shadow = false;
                nee_stage = 3;
            } else if (phase == PH_DONE && nee_stage == 0) {
                outcome = finish1d<PROC>(P, R, corr, observed, obs_x, obs_y, ox, oy, oz, qx, qy, qz, dx, dy, dz, normal_seed,
                                   t_end, ndv_end, hit, gxv, gyv, gzv);
                if (!PATHS) {
                    if (use_light && (outcome & 0xFFu) == GPIS_HIT) {
                        // one more evaluation at the scattering point sample.p = pos + t * dir (float)
                        phase = PH_NEE;
                        qx = mul_add_exact(t_end, dx, ox);
                        qy = mul_add_exact(t_end, dy, oy);
                        qz = mul_add_exact(t_end, dz, oz);
                    } else {
                        if (use_light) { nee.gx[ray] = nee.gy[ray] = nee.gz[ray] = 0.f; nee.pdf[ray] = 0.f; }
                        nee_stage = 2;
                    }
                } else {
                    const unsigned oc = outcome & 0xFFu;
                    atomicAdd(&bstat[ST_FLIGHTS], 1u);
                    atomicAdd(&bstat[ST_EVALS], evals_flight);
                    atomicAdd(&bstat[ST_STEPS], (unsigned)k);
                    atomicAdd(&bstat[oc == GPIS_HIT ? ST_HITS : (oc == GPIS_EXIT ? ST_EXITS : ST_FAILS)], 1u);
                    pev += evals_flight;
                    evals_flight = 0;
                    ++pfl;
                    if (pb == 0) {
                        pout.t0[ray] = t_end; pout.gx0[ray] = gxv; pout.gy0[ray] = gyv; pout.gz0[ray] = gzv;
                        pout.outcome0[ray] = outcome;
                    }
                    t_last = t_end;
                    o_last = outcome;
                    if (oc == GPIS_HIT) {
                        ++phit;
                        // finishDistance: sample.p = ray.pos() + sample.t * ray.dir() in float
                        hx = mul_add_exact(t_end, dx, ox);
                        hy = mul_add_exact(t_end, dy, oy);
                        hz = mul_add_exact(t_end, dz, oz);
                        hgx = gxv; hgy = gyv; hgz = gzv;
                        mdx = dx; mdy = dy; mdz = dz;
                        nee_stage = 3;  // the bounce, unless a light sample comes first
                        if (use_light) {
                            Frame3 Fh;
                            float wix, wiy, wiz;
                            hit_frame(hgx, hgy, hgz, mdx, mdy, mdz, Fh, wix, wiy, wiz);
                            float dlx = lx - hx, dly = ly - hy, dlz = lz - hz;
                            nee_rsq = dot3(dlx, dly, dlz, dlx, dly, dlz);
                            const float dist = sqrtf(nee_rsq);
                            dlx = dlx / dist; dly = dly / dist; dlz = dlz / dist;
                            float wox = dot3(Fh.tx, Fh.ty, Fh.tz, dlx, dly, dlz);
                            float woy = dot3(Fh.bx, Fh.by, Fh.bz, dlx, dly, dlz);
                            float woz = dot3(Fh.nx, Fh.ny, Fh.nz, dlx, dly, dlz);
                            normalize3(wox, woy, woz);
                            if (pin.phase == GPIS_PHASE_MIRROR) {
                                // samplingScheme at the endpoint (:68-76); a specular phase
                                // under UNI takes no light sample (TraceBase.cpp:806-807)
                                float sx, sy, sz;
                                ray_point(ox, oy, oz, dx, dy, dz, tp, sx, sy, sz);
                                const bool uni = P.scheme1d == GPIS_SCHEME_UNI || P.amp_const < 1e-6f ||
                                                 (P.corr_xy && observed && sx == ox && sy == oy && sz == oz);
                                if (!uni) {
                                    // MirrorBsdf::eval / evalGrad: the half-vector normal in world space
                                    const float nlx = (wix + wox) * 0.5f, nly = (wiy + woy) * 0.5f,
                                                nlz = (wiz + woz) * 0.5f;
                                    frame_to_global(Fh, nlx, nly, nlz, hnx, hny, hnz);
                                    normalize3(hnx, hny, hnz);
                                    phase = PH_NEE;
                                    nee_stage = 0;
                                    qx = hx; qy = hy; qz = hz;
                                }
                            } else if (wiz > 0.0f && woz > 0.0f) {
                                nee_f = kInvPiF * woz;  // LambertBsdf::eval (LambertBsdf.cpp:40-47)
                                start_shadow(hgx, hgy, hgz);
                            }
                        }
                    } else {
                        end_path();
                    }
                }
            } else if (phase == PH_DONE && nee_stage == 1) {
                // mirror normal reflecting the ray toward the light, then neeGrad / neePDF
                float hnx_, hny_, hnz_;
                if (PATHS) {
                    hnx_ = hnx; hny_ = hny; hnz_ = hnz;
                } else {
                    float wlx = lx - qx, wly = ly - qy, wlz = lz - qz;
                    normalize3(wlx, wly, wlz);
                    hnx_ = wlx - dx; hny_ = wly - dy; hnz_ = wlz - dz;
                    normalize3(hnx_, hny_, hnz_);
                }
                double tmx = 0.0, tmy = 0.0, tvar = 0.5;
                if (corr) transverse_moments(R, observed, obs_x, obs_y, ox, oy, oz, qx, qy, qz, tmx, tmy, tvar);
                float ngx, ngy, ngz, pdf;
                nee1d<PROC>(P, R, hnx_, hny_, hnz_, qx, qy, qz, ndv_nee, dx, dy, dz, corr, tmx, tmy, tvar, ngx, ngy, ngz,
                      pdf);
                if (!PATHS) {
                    nee.gx[ray] = ngx; nee.gy[ray] = ngy; nee.gz[ray] = ngz; nee.pdf[ray] = pdf;
                    nee_stage = 2;
                } else {
                    // the evaluation at the scattering point counts to the path
                    atomicAdd(&bstat[ST_EVALS], evals_flight);
                    pev += evals_flight;
                    evals_flight = 0;
                    nee_f = pdf;  // albedo 1 times neePDF (MirrorBsdf::eval)
                    if (nee_f != 0.0f) start_shadow(ngx, ngy, ngz);
                    else nee_stage = 3;
                }
            }
            if (!PATHS && phase == PH_DONE && nee_stage == 2) {
                const unsigned oc = outcome & 0xFFu;
                atomicAdd(&bstat[ST_FLIGHTS], 1u);
                atomicAdd(&bstat[ST_EVALS], evals_flight);
                atomicAdd(&bstat[ST_STEPS], (unsigned)k);
                atomicAdd(&bstat[oc == GPIS_HIT ? ST_HITS : (oc == GPIS_EXIT ? ST_EXITS : ST_FAILS)], 1u);
                out.t[ray] = t_end; out.value[ray] = val_end;
                out.gx[ray] = gxv; out.gy[ray] = gyv; out.gz[ray] = gzv;
                out.outcome[ray] = outcome;
                if (out.evals) out.evals[ray] = evals_flight;
                phase = PH_NEED;
            }
            if (PATHS && phase == PH_DONE && nee_stage == 3) {
                // the phase sample and the next flight of the path (BRDFPhaseFunction::sample)
                bool more = pb < pin.max_bounces;
                float wx = 0.f, wy = 0.f, wz = 0.f, jit = 0.f;
                if (more) {
                    Frame3 Fh;
                    float wix, wiy, wiz;
                    hit_frame(hgx, hgy, hgz, mdx, mdy, mdz, Fh, wix, wiy, wiz);
                    Pcg g;
                    g.set((uint64_t)(uint32_t)(xxh4(ray, (uint32_t)(pb + 1), pin.bounce_seed, 0u) + 1u));
                    const float u1 = u01(g.next()), u2 = u01(g.next());
                    jit = u01(g.next());
                    float ax, ay, az;
                    if (pin.phase == GPIS_PHASE_MIRROR) {
                        ax = -wix; ay = -wiy; az = wiz;  // MirrorBsdf::sample
                    } else {
                        // LambertBsdf::sample: SampleWarp::cosineHemisphere(next2D())
                        more = wiz > 0.0f;
                        const float ph = u1 * kTwoPiF, rr = sqrtf(u2);
                        ax = cosf(ph) * rr;
                        ay = sinf(ph) * rr;
                        az = sqrtf(fmaxf(1.0f - u2, 0.0f));
                    }
                    frame_to_global(Fh, ax, ay, az, wx, wy, wz);
                    normalize3(wx, wy, wz);
                }
                if (!more) {
                    end_path();
                } else {
                    ++pb;
                    ox = hx; oy = hy; oz = hz;
                    dx = wx; dy = wy; dz = wz;
                    near_t = 0.f;
                    far_t = box_exit(hx, hy, hz, wx, wy, wz, pin.box);
                    first = 0;
                    prev_v = 0.f; prev_gx = hgx; prev_gy = hgy; prev_gz = hgz;
                    path_seeds(P, pin, ray, w0 + (uint32_t)pb, R.seed, normal_seed);
                    // beginDistance (GaussianProcessMedium.cpp:244-265): medium bounce limit,
                    // empty interval
                    if ((int)(w0 + (uint32_t)pb) >= P.max_bounce) no_march_flight(GPIS_FAIL | (GPIS_FAIL_BOUNCE_LIMIT << 8));
                    else if (far_t <= near_t) no_march_flight(GPIS_EXIT);
                    else begin_flight(jit);
                }
            }
            const unsigned long long need = __ballot(phase == PH_NEED);
            if (need) {
                unsigned cnt = (unsigned)__popcll(need);
                unsigned leader = (unsigned)(__ffsll((long long)need) - 1);
                unsigned base = 0;
                if (lane == leader) base = atomicAdd(work, cnt);
                base = __shfl(base, leader);
                if (phase == PH_NEED) {
                    const unsigned rl = base + mask_rank(need);
                    if (rl >= nn) {
                        phase = PH_DEAD;
                    } else {
                        const unsigned r = idx ? idx[rl] : rl;
                        ray = r;
                        ox = in.ox[r]; oy = in.oy[r]; oz = in.oz[r];
                        dx = in.dx[r]; dy = in.dy[r]; dz = in.dz[r];
                        near_t = in.near_t[r]; far_t = in.far_t[r];
                        float jit = in.jitter[r];
                        first = (in.flags[r] & 1u) ? 1 : 0;
                        if (!first) { prev_v = in.pv[r]; prev_gx = in.pgx[r]; prev_gy = in.pgy[r]; prev_gz = in.pgz[r]; }
                        if (PATHS) {
                            w0 = pin.w0[r];
                            pb = 0;
                            shadow = false;
                            pfl = phit = psh = pvis = pev = 0;
                            nee_acc = 0.f;
                            path_seeds(P, pin, r, w0, R.seed, normal_seed);
                        } else {
                            R.seed = in.noise_seed[r];
                            normal_seed = in.normal_seed[r];
                        }
                        ray_scales(P.aniso, P.l_aniso, P.l_aniso_inv, lscale, r, R.la, R.ia);
                        if (!isfinite(far_t)) far_t = near_t + 2000.f;
                        bool bad = !isfinite(ox) || !isfinite(oy) || !isfinite(oz) || !isfinite(dx) || !isfinite(dy) ||
                                   !isfinite(dz) || !isfinite(near_t) || !(jit >= 0.f && jit < 1.f) ||
                                   !(R.la[0] > 0.f && R.la[1] > 0.f && R.la[2] > 0.f) || !isfinite(R.la[0] + R.la[1] + R.la[2]);
                        if (PATHS && bad) {
                            pout.t0[r] = far_t; pout.gx0[r] = pout.gy0[r] = pout.gz0[r] = 0.f;
                            pout.outcome0[r] = GPIS_INVALID;
                            t_last = far_t;
                            o_last = GPIS_INVALID;
                            atomicAdd(&bstat[ST_INVALID], 1u);
                            end_path();
                        } else if (PATHS && (int)w0 >= P.max_bounce) {
                            no_march_flight(GPIS_FAIL | (GPIS_FAIL_BOUNCE_LIMIT << 8));
                        } else if (PATHS && far_t <= near_t) {
                            no_march_flight(GPIS_EXIT);
                        } else if (bad || far_t <= near_t) {
                            out.t[r] = far_t; out.value[r] = (bad || first) ? 0.f : prev_v;
                            out.gx[r] = out.gy[r] = out.gz[r] = 0.f;
                            out.outcome[r] = bad ? GPIS_INVALID : GPIS_EXIT;
                            if (out.evals) out.evals[r] = 0;
                            if (use_light) { nee.gx[r] = nee.gy[r] = nee.gz[r] = 0.f; nee.pdf[r] = 0.f; }
                            atomicAdd(&bstat[bad ? ST_INVALID : ST_EXITS], 1u);
                            phase = PH_NEED;
                        } else {
                            begin_flight(jit);
                        }
                    }
                }
            }
        }
        if (!__ballot(phase != PH_DEAD)) break;
        if (phase == PH_NEED || phase == PH_DEAD || phase == PH_DONE) continue;

        float fv, nv, ndv;
        eval1d_shared<NI, PROC, FAST>(P, R, C, nf, qx, qy, qz, phase != PH_COND, fv, nv, ndv);
        ++evals_flight;
        double tq = 0.0;
        bool unit = false;

        if (phase == PH_COND) {
            condition1d<PROC>(P, R, ox, oy, oz, nv, ndv, prev_v, prev_gx, prev_gy, prev_gz, bx1, by1, obs_x, obs_y, observed);
            phase = PH_F0;
            tq = (double)near_t;
        } else if (phase == PH_F0) {
            sign0 = fv < 0.f ? -1 : 1;
            pf = fv;
            k = 0;
            if (t0 < far_t) {
                phase = PH_MARCH;
                tq = (double)t0;
            } else {
                phase = PH_FAR;
                tq = (double)far_t;
            }
        } else if (phase == PH_MARCH) {
            ++k;
            int sc = fv < 0.f ? -1 : 1;
            tk = lattice_t(t0, k - 1, step);
            if (!first && k == 1) {
                sign0 = sc;
            } else if (sc != sign0) {
                fac = (double)pf / ((double)pf - (double)fv);
                tp = lerp_d(tk - (double)step, tk, fac);
                tt = tp;
                phase = PH_SECANT;
                tq = tt;
            }
            if (phase == PH_MARCH) {
                pf = fv;
                double tn = lattice_t(t0, k, step);
                if (k >= kMaxSteps) {
                    atomicAdd(&bstat[ST_INVALID], 1u);
                    if (!PATHS) {
                        out.t[ray] = (float)tn; out.value[ray] = fv;
                        out.gx[ray] = out.gy[ray] = out.gz[ray] = 0.f;
                        out.outcome[ray] = GPIS_INVALID;
                        if (out.evals) out.evals[ray] = evals_flight;
                        if (use_light) { nee.gx[ray] = nee.gy[ray] = nee.gz[ray] = 0.f; nee.pdf[ray] = 0.f; }
                        phase = PH_NEED;
                    } else {
                        // step limit (RaySegmentStatus::StepLimit): the path ends invalid
                        if (!shadow) {
                            if (pb == 0) {
                                pout.t0[ray] = (float)tn; pout.gx0[ray] = pout.gy0[ray] = pout.gz0[ray] = 0.f;
                                pout.outcome0[ray] = GPIS_INVALID;
                            }
                            ++pfl;
                        }
                        pev += evals_flight;
                        t_last = (float)tn;
                        o_last = GPIS_INVALID;
                        end_path();
                    }
                } else if (tn < (double)far_t) {
                    tq = tn;
                } else {
                    phase = PH_FAR;
                    tq = (double)far_t;
                }
            }
        } else if (phase == PH_SECANT) {
            bool done = false;
            if ((fv < 0.f ? -1 : 1) == sign0) {
                done = true;
            } else {
                fac *= 0.9;
                if (fac <= 0.01) {
                    tp = 0.0;
                    done = true;
                } else {
                    tp = tt;
                    tt = lerp_d(tk - (double)step, tk, fac);
                    tq = tt;
                }
            }
            if (done) {
                hit = 1;
                t_end = (float)tp;
                val_end = 0.f;
                // a shadow flight needs only whether it crossed
                phase = (PATHS && shadow) ? PH_DONE : PH_GRAD;
                tq = tp;
                unit = true;
            }
        } else if (phase == PH_FAR) {
            hit = 0;
            t_end = far_t;
            val_end = fv;
            phase = (PATHS && shadow) ? PH_DONE : PH_GRAD;
            tq = (double)far_t;
            unit = true;
        } else if (phase == PH_GRAD) {
            ndv_end = ndv;      // the gradient itself is formed in the batched finish step
            phase = PH_DONE;
        } else if (phase == PH_NEE) {
            ndv_nee = ndv;
            nee_stage = 1;
            phase = PH_DONE;
        }
        if (phase == PH_F0 || phase == PH_MARCH || phase == PH_SECANT || phase == PH_FAR || phase == PH_GRAD)
            point_at(ox, oy, oz, dx, dy, dz, tq, unit, qx, qy, qz);
    }
    __syncthreads();
    if (threadIdx.x < ST_N && bstat[threadIdx.x]) atomicAdd(&stats[threadIdx.x], (unsigned long long)bstat[threadIdx.x]);
}

// ------------------------------------------------- the endpoint of a served flight --
// The learned sampler returns a distance, not a realization.  resumeDistance still draws a
// gradient at the endpoint and applies its rejection rules (GaussianProcessMedium.cpp:300-330),
// and an exit carries the field value at the bound (the next flight's conditioning value,
// SparseConvolutionNoiseMedium.cpp:339-346).  A served endpoint takes them from the law of
// the flight's conditioned 1D SE process psi(u) (isotropic ray units u = k s, k =
// |worldToLocal(d)|, correlation exp(-u^2/4)), f(s) = m(s) + sigma psi(k s), observed at the
// origin as psi(0) = a and, for Renewal+, psi'(0) = b (conditioning1D,
// SparseConvolutionNoise.cpp:576-656):
//   * the crossing's sense is the sign of f at the first lattice point t0 = near + step *
//     jitter, where a conditioned flight takes its sign reference (:304-306), drawn from its
//     conditional law;
//   * at a crossing (f(t) = 0) the derivative along the ray has the Rice density
//     |y| N(y; mu, s^2) on the sense's side (y < 0: a down-crossing, a hit; y > 0: the
//     outward crossing the rules reject);
//   * at the t = 0 atom (the secant's cut-off returns the ray origin) the derivative is the
//     conditioned one there (b under Renewal+, N(0, 1/2) under Renewal);
//   * at an exit the value follows the conditional law of f(far) on the sense's side and
//     the derivative its law given that value.
// The transverse gradient is the plain method's own Box-Muller draw (finish1d), so the
// served gradient differs from a plain one of the same ray only along the ray.  The law's
// draws use the normal seed's PCG stream after the four draws of the plain method's pairs.

// Conditional mean and covariance of (psi(x), psi'(x)) at lag x (isotropic units) given
// psi(0) = a and, when `plus`, psi'(0) = b.
struct Law1D { double e1, e2, c11, c12, c22; };
__device__ __forceinline__ Law1D cond_law1d(double dl, double a, double b, bool plus) {
    const double x = 0.5 * dl * dl;
    const double rho = exp(-0.5 * x), r2 = rho * rho;
    Law1D L;
    if (plus) {
        L.e1 = rho * (a + dl * b);
        L.e2 = rho * (-0.5 * dl * a + (1.0 - x) * b);
        // 1 - e^{-x} (1 + x) by its series where the difference cancels
        L.c11 = x < 1e-3 ? x * x * (0.5 + x * (-1.0 / 3.0 + x * (1.0 / 8.0 + x * (-1.0 / 30.0 + x / 144.0))))
                         : 1.0 - r2 * (1.0 + x);
        L.c22 = 0.5 * (1.0 - r2 * (1.0 - x + x * x));
        L.c12 = 0.25 * dl * dl * dl * r2;
    } else {
        L.e1 = rho * a;
        L.e2 = -0.5 * dl * rho * a;
        L.c11 = -expm1(-x);
        L.c22 = 0.5 * (1.0 - x * r2);
        L.c12 = 0.5 * dl * r2;
    }
    return L;
}

// Draws of the endpoint law: uniforms of a PCG stream and Box-Muller normals (FP32).
struct LawDraws {
    Pcg g;
    __device__ __forceinline__ float uni() { return u01(g.next()); }
    __device__ __forceinline__ float normal() {
        const float u1 = uni(), u2 = uni();
        return sqrtf(-2.0f * __logf(1.0f - u1)) * __cosf(6.28318530718f * u2);
    }
};

// X >= 0 with density proportional to x exp(-x^2/2 + m x), the size-biased normal N(m, 1)
// on [0, inf), by rejection: for m < 0 from a Gamma(2) proposal of rate 1/mode, accepted with
// exp(-(x - mode)^2 / 2); for m >= 0 from a unit normal about m + 1/mode, accepted with
// (x/mode) exp(1 - x/mode).  Both acceptances are at least about 0.4; one round draws both
// proposals so the lanes of a wave stay on one path whatever their m.
__device__ __forceinline__ float rice_draw(float m, LawDraws &D) {
    m = fminf(fmaxf(m, -1e6f), 1e6f);
    const float sq = sqrtf(m * m + 4.0f);
    const float xm = m < 0.f ? 2.0f / (sq - m) : 0.5f * (m + sq);
    const float inv_xm = 1.0f / xm;
    const bool gam = m < 0.f;
    for (int it = 0; it < 64; ++it) {
        const float u1 = D.uni(), u2 = D.uni(), u3 = D.uni();
        // Gamma(2): -log(u u') * mode; normal: Box-Muller of (u1, u2)
        const float lg = __logf((1.0f - u1) * (1.0f - u2));
        const float xn = m + inv_xm + sqrtf(-2.0f * __logf(1.0f - u1)) * __cosf(6.28318530718f * u2);
        const float x = gam ? -lg * xm : xn;
        const float r = x * inv_xm;
        const float acc = gam ? __expf(-0.5f * (x - xm) * (x - xm)) : r * __expf(1.0f - r);
        if (x > 0.f && u3 < acc) return x;
    }
    return xm;
}

// Z ~ N(0, 1) restricted to z > alpha (FP32): the inverse CDF for alpha < 3, Robert's
// exponential rejection above.
__device__ __forceinline__ float normal_above(float alpha, LawDraws &D) {
    alpha = fminf(alpha, 1e6f);
    if (alpha < 3.0f) {
        const float q = 0.5f * erfcf(alpha * 0.70710678f);  // P(Z > alpha)
        const float u = 1.0f - D.uni();                      // (0, 1]
        return 1.41421356f * erfcinvf(2.0f * u * q);
    }
    for (int it = 0; it < 64; ++it) {
        const float z = alpha - __logf(1.0f - D.uni()) / alpha;
        const float dz = z - alpha;
        if (D.uni() < __expf(-0.5f * dz * dz)) return z;
    }
    return alpha + 1.0f / alpha;
}

// The endpoint of served row r (its outcome word holds the pending bit and the class).
template <bool PROC>
__device__ __forceinline__ void served_finish_row(const FieldParams &P, const RayIn &in, const float *__restrict__ lscale,
                                               unsigned r, const ReplyOut &out) {
#pragma clang fp contract(off)
    const uint32_t code = out.outcome[r];
    if (!(code & mlp::kServedPending)) return;
    const int cls = (int)(code & 0xFFu);
    const float ox = in.ox[r], oy = in.oy[r], oz = in.oz[r];
    const float dx = in.dx[r], dy = in.dy[r], dz = in.dz[r];
    const float near_t = in.near_t[r], far_t = in.far_t[r];
    const uint32_t normal_seed = in.normal_seed[r];
    Ray1D R = {};
    ray_scales(P.aniso, P.l_aniso, P.l_aniso_inv, lscale, r, R.la, R.ia);
    float ix = R.ia[0] * dx, iy = R.ia[1] * dy, iz = R.ia[2] * dz;
    const double k = sqrt((double)ix * ix + (double)iy * iy + (double)iz * iz);  // du/ds
    normalize3(ix, iy, iz);
    R.F = make_frame(ix, iy, iz);
    R.cond = 1;
    const double sig = P.amp_const;
    const bool plus = P.ctxt == GPIS_CTX_RENEWAL_PLUS;
    const bool corr = P.corr_xy && plus;
    // the observations at the origin (conditioning1D's targets)
    float mgx, mgy, mgz;
    mean_grad_d<PROC>(P, ox, oy, oz, mgx, mgy, mgz);
    const double a = ((double)in.pv[r] - mean_value_d<PROC>(P, ox, oy, oz)) / sig;
    const float amp = P.amp_const;
    const float tgx = (in.pgx[r] - mgx) / amp, tgy = (in.pgy[r] - mgy) / amp, tgz = (in.pgz[r] - mgz) / amp;
    const float gwx = R.la[0] * tgx, gwy = R.la[1] * tgy, gwz = R.la[2] * tgz;
    const double b = plus ? (double)dot3(R.F.nx, R.F.ny, R.F.nz, gwx, gwy, gwz) : 0.0;
    float obs_x = 0.f, obs_y = 0.f;
    if (plus) {
        obs_x = dot3(R.F.tx, R.F.ty, R.F.tz, gwx, gwy, gwz);
        obs_y = dot3(R.F.bx, R.F.by, R.F.bz, gwx, gwy, gwz);
    }
    LawDraws D;
    D.g.set((uint64_t)normal_seed);
    for (int i = 0; i < 4; ++i) (void)D.g.next();  // past the plain method's two pairs
    // the sense: the sign of f at the first lattice point
    float step = (far_t - near_t) / P.min_step_f;
    if (P.step_size < step) step = P.step_size;
    const float t0 = mul_add_exact(step, in.jitter[r], near_t);
    float qx, qy, qz;
    ray_point(ox, oy, oz, dx, dy, dz, (double)t0, qx, qy, qz);
    const Law1D L0 = cond_law1d(k * (double)t0, a, b, plus);
    const double f0 = mean_value_d<PROC>(P, qx, qy, qz) + sig * (L0.e1 + sqrt(fmax(L0.c11, 0.0)) * (double)D.normal());
    const bool down = !(f0 < 0.0);  // a positive side looks for a down-crossing
    float t_end, ndv, value = 0.f;
    int hit = 1;
    if (cls == 0) {
        // the secant's cut-off: the ray origin and the conditioned derivative there
        t_end = 0.f;
        const Law1D L = cond_law1d(0.0, a, b, plus);
        ndv = (float)(L.e2 + sqrt(fmax(L.c22, 0.0)) * (double)D.normal());
    } else {
        t_end = cls >= mlp::kBins ? far_t : out.t[r];
        hit = cls < mlp::kBins;
    }
    grad_point(ox, oy, oz, dx, dy, dz, (double)t_end, qx, qy, qz);
    if (cls != 0) {
        const double m = mean_value_d<PROC>(P, qx, qy, qz);
        float ex, ey, ez;
        mean_grad_d<PROC>(P, qx, qy, qz, ex, ey, ez);
        const double ms = (double)ex * dx + (double)ey * dy + (double)ez * dz;  // mean slope
        const Law1D L = cond_law1d(k * (double)t_end, a, b, plus);
        const bool ok = L.c11 > 1e-300;
        const double q = ok ? L.c12 / L.c11 : 0.0;
        const double vs = fmax(L.c22 - q * L.c12, 0.0);
        const double sk = sig * k;
        if (hit) {
            // the Rice law of y = f'(t) at f(t) = 0, on the sense's side
            const double mu = L.e2 + (ok ? q * (-m / sig - L.e1) : 0.0);
            const double muy = ms + sk * mu, sy = sk * sqrt(vs);
            double y;
            if (sy > 1e-30 * (fabs(muy) + 1.0)) {
                const double w = (double)rice_draw((float)((down ? -muy : muy) / sy), D) * sy;
                y = down ? -w : w;
            } else {
                y = muy;
            }
            ndv = (float)((y - ms) / sk);
        } else {
            // the value at the bound on the sense's side, then the derivative given it
            const double sd = sqrt(fmax(L.c11, 0.0));
            double psi = L.e1;
            if (sd > 0.0) {
                const double alpha = (-m / sig - L.e1) / sd;  // psi > x0 keeps f > 0
                psi = L.e1 + sd * (double)(down ? normal_above((float)alpha, D) : -normal_above((float)-alpha, D));
            }
            value = (float)(m + sig * psi);
            const double mu = L.e2 + (ok ? q * (psi - L.e1) : 0.0);
            ndv = (float)(mu + sqrt(vs) * (double)D.normal());
        }
    }
    float gx, gy, gz;
    const uint32_t oc = finish1d<PROC>(P, R, corr, plus, obs_x, obs_y, ox, oy, oz, qx, qy, qz, dx, dy, dz,
                                       normal_seed, t_end, ndv, hit, gx, gy, gz);
    out.t[r] = t_end;
    out.value[r] = value;
    out.gx[r] = gx;
    out.gy[r] = gy;
    out.gz[r] = gz;
    out.outcome[r] = oc;
}

template <bool PROC>
__global__ __launch_bounds__(256) void served_finish_kernel(const FieldParams *__restrict__ Pg, RayIn in,
                                                            const float *__restrict__ lscale, unsigned n, ReplyOut out) {
    const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) served_finish_row<PROC>(*Pg, in, lscale, r, out);
}

// Flights of a short index list (the learned sampler's fallback rays): a group of 16 lanes
// (four groups per wave) traces one flight.  The march evaluates 16 consecutive lattice
// points at once, one per lane; conditioning, f(nearT), the secant, f(farT) and the endpoint
// gradient evaluate one point on all 16 lanes.  The points, their order of decisions and
// the arithmetic are flight1d_kernel's, so every reply is the same bit for bit (evals
// reports the plain march's count), while a long flight takes about steps/16 evaluation
// rounds: a launch over a few thousand rays no longer waits ~steps rounds for its longest
// flight.  Ray indices come from idx[0 .. *n_dev).
#ifndef GPIS_FALLBACK_GROUP
#define GPIS_FALLBACK_GROUP 16
#endif
constexpr int kGroup = GPIS_FALLBACK_GROUP;  // lanes per flight: 16, 32 or 64
static_assert(kGroup == 16 || kGroup == 32 || kGroup == 64, "a group is a power-of-two part of a wave");
constexpr unsigned long long kGroupMask = kGroup == 64 ? ~0ull : ((1ull << kGroup) - 1ull);
template <int NI, bool PROC, bool FAST>
__global__ __launch_bounds__(256, 2) void flight1d_group_kernel(const FieldParams *__restrict__ Pg, RayIn in,
                                                                const float *__restrict__ lscale, ReplyOut out,
                                                                unsigned int *work, unsigned long long *stats,
                                                                float nf, const uint32_t *__restrict__ idx,
                                                                const uint32_t *__restrict__ n_dev) {
    const FieldParams &P = *Pg;
    const unsigned nn = *n_dev;
    __shared__ unsigned int bstat[ST_N];
    if (threadIdx.x < ST_N) bstat[threadIdx.x] = 0;
    __syncthreads();
    const unsigned lane = lane_id();
    const unsigned sub = lane % kGroup;
    const unsigned gshift = lane - sub;  // first lane of this lane's group
    const unsigned long long below = gshift ? ((1ull << gshift) - 1ull) : 0ull;
    int phase = PH_NEED;
    unsigned ray = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, near_t = 0, far_t = 0, step = 0, t0 = 0;
    float pf = 0.f, t_end = 0.f, val_end = 0.f;
    int first = 1, sign0 = 1, hit = 0;
    unsigned kb = 0, k = 0, evals_u = 0, evals_m = 0;
    double tk = 0, fac = 0, tp = 0, tt = 0, tq = 0;
    bool unit = false;
    float prev_v = 0, prev_gx = 0, prev_gy = 0, prev_gz = 0;
    uint32_t normal_seed = 0;
    float bx1 = 0.f, by1 = 0.f, obs_x = 0.f, obs_y = 0.f;
    bool observed = false;
    const bool corr = P.corr_xy && P.ctxt == GPIS_CTX_RENEWAL_PLUS;
    Ray1D R = {};
    Cells1D<NI> C;
    C.base = 0;
    bool cache_valid = false;

    for (;;) {
        // ---- refill: every group that wants a ray takes the next index (one atomic per wave)
        const unsigned long long need = __ballot(phase == PH_NEED && sub == 0);
        if (need) {
            const unsigned cnt = (unsigned)__popcll(need);
            const unsigned leader = (unsigned)(__ffsll((long long)need) - 1);
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(work, cnt);
            base = __shfl(base, leader);
            if (phase == PH_NEED) {
                const unsigned rl = base + (unsigned)__popcll(need & below);
                if (rl >= nn) {
                    phase = PH_DEAD;
                } else {
                    const unsigned r = idx[rl];
                    ray = r;
                    ox = in.ox[r]; oy = in.oy[r]; oz = in.oz[r];
                    dx = in.dx[r]; dy = in.dy[r]; dz = in.dz[r];
                    near_t = in.near_t[r]; far_t = in.far_t[r];
                    const float jit = in.jitter[r];
                    R.seed = in.noise_seed[r];
                    normal_seed = in.normal_seed[r];
                    R.cond = 0;
                    first = (in.flags[r] & 1u) ? 1 : 0;
                    if (!first) { prev_v = in.pv[r]; prev_gx = in.pgx[r]; prev_gy = in.pgy[r]; prev_gz = in.pgz[r]; }
                    evals_u = 0; evals_m = 0; k = 0;
                    observed = false;
                    cache_valid = false;
                    ray_scales(P.aniso, P.l_aniso, P.l_aniso_inv, lscale, r, R.la, R.ia);
                    float ix = R.ia[0] * dx, iy = R.ia[1] * dy, iz = R.ia[2] * dz;
                    normalize3(ix, iy, iz);
                    R.F = make_frame(ix, iy, iz);
                    if (!isfinite(far_t)) far_t = near_t + 2000.f;
                    const bool bad = !isfinite(ox) || !isfinite(oy) || !isfinite(oz) || !isfinite(dx) ||
                                     !isfinite(dy) || !isfinite(dz) || !isfinite(near_t) || !(jit >= 0.f && jit < 1.f) ||
                                     !(R.la[0] > 0.f && R.la[1] > 0.f && R.la[2] > 0.f) ||
                                     !isfinite(R.la[0] + R.la[1] + R.la[2]);
                    if (bad || far_t <= near_t) {
                        if (sub == 0) {
                            out.t[r] = far_t; out.value[r] = (bad || first) ? 0.f : prev_v;
                            out.gx[r] = out.gy[r] = out.gz[r] = 0.f;
                            out.outcome[r] = bad ? GPIS_INVALID : GPIS_EXIT;
                            if (out.evals) out.evals[r] = 0;
                            atomicAdd(&bstat[bad ? ST_INVALID : ST_EXITS], 1u);
                        }
                        phase = PH_NEED;
                    } else {
                        // finishGpisSegment step rule (SparseConvolutionNoiseMedium.cpp:289-291)
                        step = (far_t - near_t) / P.min_step_f;
                        if (P.step_size < step) step = P.step_size;
                        t0 = mul_add_exact(step, jit, near_t);
                        unit = false;
                        if (P.activate && !first) {
                            phase = PH_COND;
                            if (corr) {
                                double z0, z1;
                                normal_pair(normal_seed, 0, z0, z1);
                                const double sd = sqrt(0.5);
                                bx1 = (float)(0.0 + sd * z0);
                                by1 = (float)(0.0 + sd * z1);
                            } else if (P.ctxt == GPIS_CTX_RENEWAL_PLUS) {
                                double x1, y1;
                                transverse_pair(normal_seed, 0, x1, y1);
                                bx1 = (float)x1; by1 = (float)y1;
                            }
                        } else {
                            phase = PH_F0;
                            tq = (double)near_t;
                        }
                    }
                }
            }
        }
        if (!__ballot(phase != PH_DEAD)) break;
        if (phase == PH_NEED || phase == PH_DEAD) continue;

        // ---- one evaluation per lane: its own lattice point in the march, else the group's point
        float qx, qy, qz;
        if (phase == PH_COND) {
            qx = ox; qy = oy; qz = oz;
        } else if (phase == PH_MARCH) {
            point_at(ox, oy, oz, dx, dy, dz, lattice_t(t0, (int)(kb + sub), step), false, qx, qy, qz);
        } else {
            point_at(ox, oy, oz, dx, dy, dz, tq, unit, qx, qy, qz);
        }
        float fv, nv, ndv;
        eval1d<NI, PROC, FAST>(P, R, C, cache_valid, nf, qx, qy, qz, phase != PH_COND, fv, nv, ndv);

        if (phase == PH_COND) {
            ++evals_u;
            condition1d<PROC>(P, R, ox, oy, oz, nv, ndv, prev_v, prev_gx, prev_gy, prev_gz, bx1, by1, obs_x, obs_y, observed);
            phase = PH_F0;
            tq = (double)near_t;
            unit = false;
        } else if (phase == PH_F0) {
            ++evals_u;
            sign0 = fv < 0.f ? -1 : 1;
            pf = fv;
            kb = 0;
            if (t0 < far_t) {
                phase = PH_MARCH;
            } else {
                phase = PH_FAR;
                tq = (double)far_t;
            }
        } else if (phase == PH_MARCH) {
            // lattice point kb + sub; it exists while t < farT (the plain march's loop bound)
            const double tl = lattice_t(t0, (int)(kb + sub), step);
            const bool valid = tl < (double)far_t && kb + sub < (unsigned)kMaxSteps;
            const int sc = fv < 0.f ? -1 : 1;
            // a conditioned flight's first march point only sets the reference sign
            if (!first && kb == 0) sign0 = __shfl(sc, (int)gshift);
            const bool cand = valid && sc != sign0 && (first || kb + sub != 0);
            const unsigned long long cm = (__ballot(cand) >> gshift) & kGroupMask;
            const unsigned long long vm = (__ballot(valid) >> gshift) & kGroupMask;
            const float f_last = __shfl(fv, (int)(gshift + kGroup - 1));
            if (cm) {
                const unsigned j = (unsigned)(__ffsll((long long)cm) - 1);
                const float fj = __shfl(fv, (int)(gshift + j));
                const float fjm1 = __shfl(fv, (int)(gshift + (j > 0 ? j - 1 : 0)));
                const float pfx = j > 0 ? fjm1 : pf;
                const unsigned kc = kb + j;  // lattice index of the first sign change
                evals_m = kc + 1;
                k = kc + 1;
                // secant with 0.9 shrink (SparseConvolutionNoiseMedium.cpp:309-329), in FP64
                tk = lattice_t(t0, (int)kc, step);
                fac = (double)pfx / ((double)pfx - (double)fj);
                tp = lerp_d(tk - (double)step, tk, fac);
                tt = tp;
                phase = PH_SECANT;
                tq = tt;
                unit = false;
            } else if (vm != kGroupMask) {
                evals_m = kb + (unsigned)__popcll(vm);
                k = evals_m;
                phase = PH_FAR;
                tq = (double)far_t;
                unit = false;
            } else {
                pf = f_last;
                kb += kGroup;
                if (kb >= (unsigned)kMaxSteps) {
                    // step limit (RaySegmentStatus::StepLimit): reported invalid, as the plain kernel
                    if (sub == 0) {
                        out.t[ray] = (float)lattice_t(t0, kMaxSteps, step); out.value[ray] = f_last;
                        out.gx[ray] = out.gy[ray] = out.gz[ray] = 0.f;
                        out.outcome[ray] = GPIS_INVALID;
                        if (out.evals) out.evals[ray] = evals_u + (unsigned)kMaxSteps;
                        atomicAdd(&bstat[ST_INVALID], 1u);
                    }
                    phase = PH_NEED;
                }
            }
        } else if (phase == PH_SECANT) {
            ++evals_u;
            bool done = false;
            if ((fv < 0.f ? -1 : 1) == sign0) {
                done = true;
            } else {
                fac *= 0.9;
                if (fac <= 0.01) {
                    tp = 0.0;
                    done = true;
                } else {
                    tp = tt;
                    tt = lerp_d(tk - (double)step, tk, fac);
                    tq = tt;
                }
            }
            if (done) {
                hit = 1;
                t_end = (float)tp;
                val_end = 0.f;
                phase = PH_GRAD;
                tq = tp;
                unit = true;
            }
        } else if (phase == PH_FAR) {
            ++evals_u;
            hit = 0;
            t_end = far_t;
            val_end = fv;
            phase = PH_GRAD;
            tq = (double)far_t;
            unit = true;
        } else if (phase == PH_GRAD) {
            ++evals_u;
            float gxv, gyv, gzv;
            const uint32_t outcome = finish1d<PROC>(P, R, corr, observed, obs_x, obs_y, ox, oy, oz, qx, qy, qz, dx, dy, dz,
                                              normal_seed, t_end, ndv, hit, gxv, gyv, gzv);
            if (sub == 0) {
                const unsigned oc = outcome & 0xFFu;
                atomicAdd(&bstat[ST_FLIGHTS], 1u);
                atomicAdd(&bstat[ST_EVALS], evals_u + evals_m);
                atomicAdd(&bstat[ST_STEPS], k);
                atomicAdd(&bstat[oc == GPIS_HIT ? ST_HITS : (oc == GPIS_EXIT ? ST_EXITS : ST_FAILS)], 1u);
                out.t[ray] = t_end; out.value[ray] = val_end;
                out.gx[ray] = gxv; out.gy[ray] = gyv; out.gz[ray] = gzv;
                out.outcome[ray] = outcome;
                if (out.evals) out.evals[ray] = evals_u + evals_m;
            }
            phase = PH_NEED;
        }
    }
    __syncthreads();
    if (threadIdx.x < ST_N && bstat[threadIdx.x]) atomicAdd(&stats[threadIdx.x], (unsigned long long)bstat[threadIdx.x]);
}

template <bool PATHS>
static void launch_flights(gpis_ctx *c, unsigned blocks, unsigned threads, hipStream_t st, RayIn in, ReplyOut out,
                           PathOut pout, PathSeeds3D pseeds, unsigned n, int max_bounces, uint32_t bounce_seed,
                           float box) {
    // the wave-cooperative cells' LDS: one part per wave of the block (cells_coop)
    const size_t coop_lds = c->P.coop_cells ? (size_t)((threads + 63) / 64) * coop_wave_slots(c->P.N) * sizeof(float4) : 0;
#define GPIS_FLIGHT_LAUNCH(NI, ISO, COND)                                                                          \
    do {                                                                                                           \
        if (grid)                                                                                                  \
            hipLaunchKernelGGL((flight_kernel<NI, ISO, COND, PATHS, true>), dim3(blocks), dim3(threads), coop_lds, \
                               st, c->P_dev, in, out, pout, pseeds, n, c->work, c->counters, max_bounces,          \
                               bounce_seed, box, c->batch_k);                                                      \
        else                                                                                                       \
            hipLaunchKernelGGL((flight_kernel<NI, ISO, COND, PATHS, false>), dim3(blocks), dim3(threads),          \
                               coop_lds, st, c->P_dev, in, out, pout, pseeds, n, c->work, c->counters,             \
                               max_bounces, bounce_seed, box, c->batch_k);                                         \
    } while (0)
    const bool cond = c->P.activate != 0;
    const bool grid = c->P.mean_kind == GPIS_MEAN_GRID || c->P.mean_kind == GPIS_MEAN_PROCEDURAL;
    if (c->P.iso3d) {
        // the isotropic ray space: generic impulse count, isotropic kernel, analytic means
        if (cond)
            hipLaunchKernelGGL((flight_kernel<0, true, true, PATHS, false, true>), dim3(blocks), dim3(threads), 0, st,
                               c->P_dev, in, out, pout, pseeds, n, c->work, c->counters, max_bounces, bounce_seed, box, c->batch_k);
        else
            hipLaunchKernelGGL((flight_kernel<0, true, false, PATHS, false, true>), dim3(blocks), dim3(threads), 0, st,
                               c->P_dev, in, out, pout, pseeds, n, c->work, c->counters, max_bounces, bounce_seed, box, c->batch_k);
        return;
    }
    // "noise" variance or length-scale fields and the full anisotropy matrix: the generic
    // kernels that carry them (NI = -1)
    const bool noise = c->P.var_field == 1 || c->P.ls_field == GPIS_LS_NOISE || c->P.full;
    if (noise && c->P.iso && !cond) GPIS_FLIGHT_LAUNCH(-1, true, false);
    else if (noise && c->P.iso) GPIS_FLIGHT_LAUNCH(-1, true, true);
    else if (noise && !cond) GPIS_FLIGHT_LAUNCH(-1, false, false);
    else if (noise) GPIS_FLIGHT_LAUNCH(-1, false, true);
    else if (c->P.N == 8 && c->P.iso && !cond) GPIS_FLIGHT_LAUNCH(8, true, false);
    else if (c->P.N == 8 && c->P.iso) GPIS_FLIGHT_LAUNCH(8, true, true);
    else if (c->P.N == 8 && !cond) GPIS_FLIGHT_LAUNCH(8, false, false);
    else if (c->P.N == 8) GPIS_FLIGHT_LAUNCH(8, false, true);
    else if (c->P.iso && !cond) GPIS_FLIGHT_LAUNCH(0, true, false);
    else if (c->P.iso) GPIS_FLIGHT_LAUNCH(0, true, true);
    else if (!cond) GPIS_FLIGHT_LAUNCH(0, false, false);
    else GPIS_FLIGHT_LAUNCH(0, false, true);
#undef GPIS_FLIGHT_LAUNCH
}

// ---------------------------------------------------------------- C ABI --
extern "C" {

gpis_status gpis_create(const gpis_field_desc *desc, gpis_ctx **out) {
    if (!desc || !out) return GPIS_INVALID_ARGUMENT;
    *out = nullptr;
    gpis_ctx *c = new (std::nothrow) gpis_ctx();
    if (!c) return GPIS_OUT_OF_MEMORY;
    c->desc = *desc;
    gpis_status s = build_params(c);
    if (s != GPIS_OK) {
        fprintf(stderr, "gpis_create: %s\n", c->err.c_str());
        delete c;
        return s;
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e == hipSuccess) e = hipMalloc((void **)&c->counters, 16 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->counters, 0, 16 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&c->work, 64);
    if (e == hipSuccess && c->desc.mean_kind == GPIS_MEAN_PROCEDURAL) {
        e = hipMalloc((void **)&c->sdf_dev, sizeof(sdf::MeanParams));
        if (e == hipSuccess) e = hipMemcpy(c->sdf_dev, &c->sdf_host, sizeof(sdf::MeanParams), hipMemcpyHostToDevice);
        c->P.sdf = c->sdf_dev;
    }
    c->P.coop_cells = coop_eligible(c->P);  // the medium's parameters are complete here
    if (e == hipSuccess) e = hipMalloc((void **)&c->P_dev, sizeof(FieldParams));
    if (e == hipSuccess) e = hipMemcpy(c->P_dev, &c->P, sizeof(FieldParams), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fprintf(stderr, "gpis_create: %s\n", hipGetErrorString(e));
        if (c->counters) (void)hipFree(c->counters);
        if (c->work) (void)hipFree(c->work);
        if (c->sdf_dev) (void)hipFree(c->sdf_dev);
        if (c->P_dev) (void)hipFree(c->P_dev);
        delete c;
        return GPIS_DEVICE_ERROR;
    }
    *out = c;
    return GPIS_OK;
}

gpis_status gpis_destroy(gpis_ctx *c) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    (void)hipDeviceSynchronize();
    free_lists(c);
    if (c->table) (void)hipFree(c->table);
    if (c->grid) (void)hipFree(c->grid);
    if (c->sdf_dev) (void)hipFree(c->sdf_dev);
    if (c->counters) (void)hipFree(c->counters);
    if (c->work) (void)hipFree(c->work);
    if (c->P_dev) (void)hipFree(c->P_dev);
    delete c;
    return GPIS_OK;
}

const char *gpis_last_error(const gpis_ctx *c) { return c ? c->err.c_str() : "null context"; }

static gpis_status gpis_set_mean_grid_impl(gpis_ctx *c, const gpis_mean_grid *g) {
    if (!c || !g || !g->voxels) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    if (c->P.mean_kind != GPIS_MEAN_GRID) return fail(c, GPIS_INVALID_ARGUMENT, "the medium's mean is not tabulated");
    size_t count = 1;
    for (int i = 0; i < 3; ++i) {
        // the same shape rule as the reference's dense mean grid (RaySegmentBatch.cu:312-313)
        if (g->dims[i] < 2) return fail(c, GPIS_INVALID_ARGUMENT, "grid dimensions must be at least 2");
        if (!(g->clamp_min[i] <= g->clamp_max[i])) return fail(c, GPIS_INVALID_ARGUMENT, "empty clamp box");
        count *= (size_t)g->dims[i];
    }
    if (count > ((size_t)1 << 34)) return fail(c, GPIS_INVALID_ARGUMENT, "grid too large");
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(g->world_to_index[i])) return fail(c, GPIS_INVALID_ARGUMENT, "non-finite transform");
    if (!std::isfinite(g->offset) || !std::isfinite(g->scale)) return fail(c, GPIS_INVALID_ARGUMENT, "non-finite offset or scale");
    HIP_TRY(c, hipDeviceSynchronize());
    if (c->grid) { HIP_TRY(c, hipFree(c->grid)); c->grid = nullptr; }
    c->P.grid = nullptr;
    HIP_TRY(c, hipMalloc((void **)&c->grid, count * sizeof(float)));
    HIP_TRY(c, hipMemcpy(c->grid, g->voxels, count * sizeof(float), hipMemcpyHostToDevice));
    FieldParams &P = c->P;
    P.grid = c->grid;
    for (int i = 0; i < 3; ++i) {
        P.gorig[i] = g->origin[i];
        P.gdim[i] = g->dims[i];
        P.gcmin[i] = g->clamp_min[i];
        P.gcmax[i] = g->clamp_max[i];
    }
    for (int i = 0; i < 12; ++i) P.w2i[i] = g->world_to_index[i];
    P.goff = g->offset;
    P.gscale = g->scale;
    P.gvolume = g->volume != 0;
    return GPIS_OK;
}

gpis_status gpis_set_mean_grid(gpis_ctx *c, const gpis_mean_grid *g) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    const gpis_status s = gpis_set_mean_grid_impl(c, g);
    const gpis_status t = sync_params(c);
    return s != GPIS_OK ? s : t;
}

static gpis_status mean_ready(gpis_ctx *c) {
    if (c->P.mean_kind == GPIS_MEAN_GRID && !c->P.grid)
        return fail(c, GPIS_INVALID_ARGUMENT, "tabulated mean: call gpis_set_mean_grid first");
    return GPIS_OK;
}

static void free_lists(gpis_ctx *c) {
    if (c->lfix) (void)hipFree(c->lfix);
    if (c->lent) (void)hipFree(c->lent);
    if (c->lrange) (void)hipFree(c->lrange);
    c->lfix = nullptr; c->lent = nullptr; c->lrange = nullptr;
    FieldParams &P = c->P;
    P.lfix = nullptr; P.lent = nullptr; P.lrange = nullptr;
    P.has_lists = 0; P.lsub = 0; P.lfix_bytes = P.lent_n = P.lrange_n = 0;
    P.lmargin = 1; P.lpair2 = 0.0f;
}

// Support lists of the table's interior cells (gpis_lists.hpp), or none when lists are off,
// there is no table, or the table is beyond the fixed-point range.  Synchronous.
static gpis_status build_lists(gpis_ctx *c) {
    HIP_TRY(c, hipDeviceSynchronize());
    free_lists(c);
    const int S = c->lists_sub;
    if (S <= 0 || !c->table) return GPIS_OK;
    const double reach = c->lists_points == 2 ? lists::pair_reach((double)c->P.step_size, (double)c->P.R) : 1.0;
    const int *lo = c->table_lo, *dim = c->table_dim;
    for (int i = 0; i < 3; ++i)
        if (lo[i] < -lists::kMaxAbsCell || lo[i] + dim[i] - 1 > lists::kMaxAbsCell) return GPIS_OK;
    const int N = c->P.N;
    const long long cells = (long long)c->table_cells;
    const long long imps = cells * N;
    if ((imps + 1) * 16 >= (1LL << 32)) return GPIS_OK;
    const long long nsub = cells * S * S * S;
    HIP_TRY(c, hipMalloc((void **)&c->lfix, (size_t)(imps + 1) * 16));
    HIP_TRY(c, hipMalloc((void **)&c->lrange, (size_t)nsub * 2 * sizeof(uint32_t)));
    const int threads = 256;
    hipLaunchKernelGGL(fix_build_kernel, dim3((unsigned)((imps + 1 + threads - 1) / threads)), dim3(threads), 0, 0,
                       c->table, lo[0], lo[1], lo[2], dim[0], dim[1], N, imps, c->lfix);
    HIP_TRY(c, hipGetLastError());
    uint32_t *groups = nullptr;
    HIP_TRY(c, hipMalloc((void **)&groups, (size_t)nsub * sizeof(uint32_t)));
    hipLaunchKernelGGL(list_count_kernel, dim3((unsigned)((nsub + threads - 1) / threads)), dim3(threads), 0, 0,
                       c->table, lo[0], lo[1], lo[2], dim[0], dim[1], dim[2], N, S, reach, nsub, groups);
    hipError_t e = hipGetLastError();
    std::vector<uint32_t> g((size_t)nsub), range((size_t)nsub * 2);
    if (e == hipSuccess) e = hipMemcpy(g.data(), groups, (size_t)nsub * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(groups);
    if (e != hipSuccess) return hip_fail(c, e, "list_count_kernel");
    unsigned long long total = 0;
    for (long long i = 0; i < nsub; ++i) {
        range[2 * i] = (uint32_t)total;
        range[2 * i + 1] = g[i];
        total += g[i];
    }
    // the kernels read the entries through a buffer resource of 32-bit byte offsets, one
    // group beyond a list's end
    if (total == 0 || (total + 2) * 16 >= (1ULL << 32)) { free_lists(c); return GPIS_OK; }
    HIP_TRY(c, hipMemcpy(c->lrange, range.data(), range.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    // the lists, then one group of the sentinel alone: the kernels read it for the groups
    // past a list's end
    HIP_TRY(c, hipMalloc((void **)&c->lent, (size_t)(total + 1) * 4 * sizeof(uint32_t)));
    const uint32_t pad[4] = {(uint32_t)(imps * 16), (uint32_t)(imps * 16), (uint32_t)(imps * 16), (uint32_t)(imps * 16)};
    HIP_TRY(c, hipMemcpy(c->lent + total * 4, pad, sizeof pad, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(list_fill_kernel, dim3((unsigned)((nsub + threads - 1) / threads)), dim3(threads), 0, 0,
                       c->table, lo[0], lo[1], lo[2], dim[0], dim[1], dim[2], N, S, reach, nsub, c->lrange, c->lent,
                       (uint32_t)(imps * 16));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipDeviceSynchronize());
    FieldParams &P = c->P;
    P.lfix = c->lfix; P.lent = c->lent; P.lrange = c->lrange;
    P.lsub = S;
    P.lfix_bytes = (uint32_t)((imps + 1) * 16);
    P.lent_n = (uint32_t)((total + 1) * 4);
    P.lrange_n = (uint32_t)nsub;
    P.lmargin = lists::list_margin(reach);
    // a second query within (reach - 1) of the first is served by the same pass (with a
    // margin below the lists' own)
    P.lpair2 = reach > 1.0 ? (float)((reach - 1.0) * (reach - 1.0) * 0.999) : 0.0f;
    P.has_lists = 1;
    return GPIS_OK;
}

static gpis_status gpis_set_domain_impl(gpis_ctx *c, const float lo[3], const float hi[3], uint32_t seed) {
    if (!c || !lo || !hi) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    if (c->P.multires) return fail(c, GPIS_UNSUPPORTED, "impulse tables cover the single-level grid");
    if (c->P.iso3d) return fail(c, GPIS_UNSUPPORTED, "the isotropic ray space has one lattice per flight direction");
    HIP_TRY(c, hipDeviceSynchronize());
    free_lists(c);
    if (c->table) { HIP_TRY(c, hipFree(c->table)); c->table = nullptr; }
    c->P.has_table = 0;
    c->P.table = nullptr;
    c->table_cells = 0;
    float R = c->P.R;
    int clo[3], dims[3];
    size_t cells = 1;
    for (int i = 0; i < 3; ++i) {
        if (!(hi[i] > lo[i])) return fail(c, GPIS_INVALID_ARGUMENT, "empty domain");
        // two cells of margin: the 3x3x3 block of every query in the box, and the 5x5x5
        // neighbourhood of the support lists that serve two queries per pass
        clo[i] = (int)floorf(lo[i] / R) - 2;
        int chi = (int)floorf(hi[i] / R) + 2;
        dims[i] = chi - clo[i] + 1;
        cells *= (size_t)dims[i];
    }
    if (cells * c->P.N * 3 * sizeof(float) > (size_t)8 << 30) return fail(c, GPIS_INVALID_ARGUMENT, "domain too large for a table");
    HIP_TRY(c, hipMalloc((void **)&c->table, cells * c->P.N * 3 * sizeof(float)));
    int threads = 256;
    unsigned blocks = (unsigned)((cells + threads - 1) / threads);
    hipLaunchKernelGGL(table_build_kernel, dim3(blocks), dim3(threads), 0, 0, clo[0], clo[1], clo[2], dims[0],
                       dims[1], dims[2], seed, c->P.N, c->table);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipDeviceSynchronize());
    c->table_cells = cells;
    c->P.table = c->table;
    for (int i = 0; i < 3; ++i) {
        c->P.tlo[i] = clo[i]; c->P.tdim[i] = dims[i];
        c->table_lo[i] = clo[i]; c->table_dim[i] = dims[i];
    }
    c->P.tseed = seed;
    c->P.table_n = (uint64_t)cells * c->P.N * 3;
    c->P.has_table = 1;
    return build_lists(c);
}

gpis_status gpis_set_domain(gpis_ctx *c, const float lo[3], const float hi[3], uint32_t seed) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    const gpis_status s = gpis_set_domain_impl(c, lo, hi, seed);
    const gpis_status t = sync_params(c);
    return s != GPIS_OK ? s : t;
}

gpis_status gpis_impulse_table(gpis_ctx *c, const int32_t lo[3], const int32_t dims[3], uint32_t seed, float *out,
                               void *stream) {
    if (!c || !lo || !dims || !out) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    long long total = 1;
    for (int i = 0; i < 3; ++i) {
        if (dims[i] <= 0) return fail(c, GPIS_INVALID_ARGUMENT, "dims must be positive");
        total *= dims[i];
    }
    if (total > (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "too many cells");
    int threads = 256;
    unsigned blocks = (unsigned)((total + threads - 1) / threads);
    hipLaunchKernelGGL(impulse_table_kernel, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, lo[0], lo[1], lo[2],
                       dims[0], dims[1], dims[2], seed, c->P.N, (float4 *)out);
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_eval_points(gpis_ctx *c, const float *x, const float *y, const float *z, int64_t n, uint32_t seed,
                             float *v, float *gx, float *gy, float *gz, void *stream) {
    if (!c || n < 0) return fail(c, GPIS_INVALID_ARGUMENT, "bad argument");
    if (n == 0) return GPIS_OK;
    if (!x || !y || !z || !v || !gx || !gy || !gz) return fail(c, GPIS_INVALID_ARGUMENT, "null pointer");
    if (gpis_status s = mean_ready(c)) return s;
    if (c->P.iso3d) return fail(c, GPIS_UNSUPPORTED, "the isotropic ray space needs a flight direction");
    int threads = 256;
    unsigned blocks = (unsigned)((n + threads - 1) / threads);
#define GPIS_EVAL_LAUNCH(NI, ISO)                                                                              \
    do {                                                                                                           \
        if (ext)                                                                                                   \
            hipLaunchKernelGGL((eval_points_kernel<NI, ISO, MEAN_EXT>), dim3(blocks), dim3(threads), 0,            \
                               (hipStream_t)stream, c->P_dev, x, y, z, (long long)n, seed, v, gx, gy, gz);            \
        else                                                                                                       \
            hipLaunchKernelGGL((eval_points_kernel<NI, ISO, MEAN_ANALYTIC>), dim3(blocks), dim3(threads), 0,       \
                               (hipStream_t)stream, c->P_dev, x, y, z, (long long)n, seed, v, gx, gy, gz);            \
    } while (0)
    const bool ext = c->P.mean_kind == GPIS_MEAN_GRID || c->P.mean_kind == GPIS_MEAN_PROCEDURAL;
    const bool noise = c->P.var_field == 1 || c->P.ls_field == GPIS_LS_NOISE || c->P.full;
    if (noise && c->P.iso) GPIS_EVAL_LAUNCH(-1, true);
    else if (noise) GPIS_EVAL_LAUNCH(-1, false);
    else if (c->P.N == 8 && c->P.iso) GPIS_EVAL_LAUNCH(8, true);
    else if (c->P.N == 8) GPIS_EVAL_LAUNCH(8, false);
    else if (c->P.iso) GPIS_EVAL_LAUNCH(0, true);
    else GPIS_EVAL_LAUNCH(0, false);
#undef GPIS_EVAL_LAUNCH
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

static gpis_status check_rays(gpis_ctx *c, const gpis_ray_soa *r, int64_t n) {
    if (c->P.one_d) return fail(c, GPIS_INVALID_ARGUMENT, "1D_sampling media use gpis_trace_flights_1d");
    if (!r) return fail(c, GPIS_INVALID_ARGUMENT, "null rays");
    if (n < 0 || n >= (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "ray count out of range");
    if (!r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz || !r->near_t || !r->far_t || !r->jitter ||
        !r->noise_seed || !r->flags)
        return fail(c, GPIS_INVALID_ARGUMENT, "null ray field");
    if (!r->prev_value || !r->prev_gx || !r->prev_gy || !r->prev_gz)
        return fail(c, GPIS_INVALID_ARGUMENT, "null previous-state field");
    if (r->length_scale) return fail(c, GPIS_UNSUPPORTED, "per-ray length scale is a 1D feature");
    return mean_ready(c);
}

static void launch_shape(gpis_ctx *c, int64_t n, unsigned &blocks, unsigned &threads) {
    threads = (unsigned)(c->block_size > 0 ? c->block_size : 256);
    int wpc = c->waves_per_cu > 0 ? c->waves_per_cu : 16;
    long long want = (long long)c->num_cus * wpc * 64 / threads;
    long long need = (n + threads - 1) / threads;
    if (want < 1) want = 1;
    blocks = (unsigned)(need < want ? need : want);
    if (blocks < 1) blocks = 1;
}

static RayIn make_in(const gpis_ray_soa *r) {
    RayIn in = {r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->near_t, r->far_t, r->jitter,
                r->noise_seed, r->normal_seed, r->flags, r->prev_value, r->prev_gx, r->prev_gy, r->prev_gz};
    return in;
}

gpis_status gpis_trace_flights(gpis_ctx *c, const gpis_ray_soa *rays, int64_t n, const gpis_reply_soa *reply,
                               void *stream) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    gpis_status s = check_rays(c, rays, n);
    if (s != GPIS_OK) return s;
    if (!reply || !reply->t || !reply->value || !reply->gx || !reply->gy || !reply->gz || !reply->outcome)
        return fail(c, GPIS_INVALID_ARGUMENT, "null reply field");
    if (n == 0) return GPIS_OK;
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(c, hipMemsetAsync(c->work, 0, sizeof(unsigned), st));
    unsigned blocks, threads;
    launch_shape(c, n, blocks, threads);
    ReplyOut out = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals};
    PathOut pout = {};
    PathSeeds3D ps = {};
    launch_flights<false>(c, blocks, threads, st, make_in(rays), out, pout, ps, (unsigned)n, 0, 0u, 0.f);
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_trace_paths(gpis_ctx *c, const gpis_ray_soa *rays, const gpis_path_seeds *seeds, int64_t n,
                             int32_t max_bounces, uint32_t bounce_seed, float box, const gpis_path_reply *reply,
                             void *stream) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    gpis_status s = check_rays(c, rays, n);
    if (s != GPIS_OK) return s;
    if (!reply || !reply->t0 || !reply->gx0 || !reply->gy0 || !reply->gz0 || !reply->outcome0 || !reply->flights ||
        !reply->hits)
        return fail(c, GPIS_INVALID_ARGUMENT, "null reply field");
    if ((reply->t_last == nullptr) != (reply->outcome_last == nullptr))
        return fail(c, GPIS_INVALID_ARGUMENT, "t_last and outcome_last are given together");
    if (seeds && (!seeds->px || !seeds->py || !seeds->spp || !seeds->bounce))
        return fail(c, GPIS_INVALID_ARGUMENT, "null seed field");
    if (max_bounces < 0 || max_bounces > 64) return fail(c, GPIS_INVALID_ARGUMENT, "max_bounces out of range");
    if (!(box > 0.f)) return fail(c, GPIS_INVALID_ARGUMENT, "box must be positive");
    if (n == 0) return GPIS_OK;
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(c, hipMemsetAsync(c->work, 0, sizeof(unsigned), st));
    unsigned blocks, threads;
    launch_shape(c, n, blocks, threads);
    ReplyOut out = {};
    PathOut pout = {reply->t0, reply->gx0, reply->gy0, reply->gz0, reply->outcome0, reply->flights, reply->hits,
                    reply->evals, reply->t_last, reply->outcome_last};
    PathSeeds3D ps = {};
    if (seeds) ps = {seeds->px, seeds->py, seeds->spp, seeds->bounce, xxh1(seeds->scene_seed), 1};
    launch_flights<true>(c, blocks, threads, st, make_in(rays), out, pout, ps, (unsigned)n, max_bounces, bounce_seed,
                         box);
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

static gpis_status gpis_set_lists_impl(gpis_ctx *c, int32_t sub, int32_t points) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (sub < 0 || sub > lists::kMaxSub) return fail(c, GPIS_INVALID_ARGUMENT, "sub-cells per cell edge must be in 0..8");
    if (points != 1 && points != 2) return fail(c, GPIS_INVALID_ARGUMENT, "points per list pass must be 1 or 2");
    c->lists_sub = sub;
    c->lists_points = points;
    c->P.pair_cells = points == 2;
    return build_lists(c);
}

gpis_status gpis_set_lists(gpis_ctx *c, int32_t sub, int32_t points) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    const gpis_status s = gpis_set_lists_impl(c, sub, points);
    const gpis_status t = sync_params(c);
    return s != GPIS_OK ? s : t;
}

gpis_status gpis_set_precision(gpis_ctx *c, int32_t mode) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (mode != GPIS_PRECISION_FAST && mode != GPIS_PRECISION_EXACT)
        return fail(c, GPIS_INVALID_ARGUMENT, "precision mode must be GPIS_PRECISION_FAST or GPIS_PRECISION_EXACT");
    c->fast1d = mode == GPIS_PRECISION_FAST;
    return GPIS_OK;
}

gpis_status gpis_set_coop_cells(gpis_ctx *c, int32_t on) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (on != 0 && on != 1) return fail(c, GPIS_INVALID_ARGUMENT, "coop_cells must be 0 or 1");
    c->P.coop_cells = on ? coop_eligible(c->P) : 0;
    return sync_params(c);
}

gpis_status gpis_set_refill(gpis_ctx *c, int32_t batch) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (batch < 0 || batch > 64) return fail(c, GPIS_INVALID_ARGUMENT, "refill batch must be in 0..64 lanes");
    c->batch_k = batch < 1 ? 1 : batch;
    c->batch_k_1d = c->batch_k;
    return GPIS_OK;
}

// The persistent 1D kernel for the medium's mean (procedural or not) and the precision
// mode of the context (FP32 fast path or the reference's exact rounding).
extern "C++" template <bool PATHS>
static void launch_flight1d_kernel(gpis_ctx *c, unsigned blocks, hipStream_t st, const gpis_ray_soa *rays, unsigned n,
                                   const ReplyOut &out, const float *light, const Nee1DOut &nee, const uint32_t *idx,
                                   const uint32_t *n_dev, const Path1DIn &pin, const Path1DOut &pout) {
    const float lx = light ? light[0] : 0.f, ly = light ? light[1] : 0.f, lz = light ? light[2] : 0.f;
#define GPIS_1D_LAUNCH(PROC, FAST)                                                                               \
    hipLaunchKernelGGL((flight1d_kernel<8, PATHS, PROC, FAST>), dim3(blocks), dim3(256), 0, st, c->P_dev,          \
                       make_in(rays), rays->length_scale, out, n, c->work, c->counters, c->nf1d, light ? 1 : 0, lx,   \
                       ly, lz, nee, idx, n_dev, c->batch_k_1d, pin, pout)
    const bool proc = c->P.mean_kind == GPIS_MEAN_PROCEDURAL;
    if (proc && c->fast1d) GPIS_1D_LAUNCH(true, true);
    else if (proc) GPIS_1D_LAUNCH(true, false);
    else if (c->fast1d) GPIS_1D_LAUNCH(false, true);
    else GPIS_1D_LAUNCH(false, false);
#undef GPIS_1D_LAUNCH
}

// Launch the persistent 1D kernel over n rays, or over the first *n_dev entries of the
// index list idx (n is then their upper bound).
static gpis_status launch_flight1d(gpis_ctx *c, const gpis_ray_soa *rays, int64_t n, const ReplyOut &out,
                                   const float *light, const Nee1DOut &nee, const uint32_t *idx,
                                   const uint32_t *n_dev, hipStream_t st) {
    HIP_TRY(c, hipMemsetAsync(c->work, 0, sizeof(unsigned), st));
    unsigned threads = 256;
    int wpc = c->waves_per_cu > 0 ? c->waves_per_cu : 8;
    long long want = (long long)c->num_cus * wpc * 64 / threads;
    long long need = (n + threads - 1) / threads;
    unsigned blocks = (unsigned)(need < want ? need : (want < 1 ? 1 : want));
    launch_flight1d_kernel<false>(c, blocks, st, rays, (unsigned)n, out, light, nee, idx, n_dev, Path1DIn{},
                                  Path1DOut{});
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_trace_paths_1d(gpis_ctx *c, const gpis_ray_soa *rays, const gpis_path_seeds *seeds, int64_t n,
                                int32_t max_bounces, int32_t phase, uint32_t bounce_seed, float box,
                                const float *light, const gpis_path1d_reply *reply, void *stream) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (!c->P.one_d) return fail(c, GPIS_INVALID_ARGUMENT, "the medium is not 1D_sampling");
    if (!rays || !seeds || !reply) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    if (n < 0 || n >= (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "path count out of range");
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->near_t ||
        !rays->far_t || !rays->jitter || !rays->flags || !rays->prev_value || !rays->prev_gx || !rays->prev_gy ||
        !rays->prev_gz)
        return fail(c, GPIS_INVALID_ARGUMENT, "null ray field");
    if (!seeds->px || !seeds->py || !seeds->spp || !seeds->bounce)
        return fail(c, GPIS_INVALID_ARGUMENT, "null seed field");
    if (!reply->t0 || !reply->gx0 || !reply->gy0 || !reply->gz0 || !reply->outcome0 || !reply->flights ||
        !reply->hits || !reply->shadows || !reply->visible || !reply->nee || !reply->t_last || !reply->outcome_last)
        return fail(c, GPIS_INVALID_ARGUMENT, "null reply field");
    if (max_bounces < 0 || max_bounces > 64) return fail(c, GPIS_INVALID_ARGUMENT, "max_bounces out of range");
    if (phase != GPIS_PHASE_MIRROR && phase != GPIS_PHASE_LAMBERT)
        return fail(c, GPIS_INVALID_ARGUMENT, "phase must be GPIS_PHASE_MIRROR or GPIS_PHASE_LAMBERT");
    if (!(box > 0.f) || !std::isfinite(box)) return fail(c, GPIS_INVALID_ARGUMENT, "box must be positive");
    if (light && !(std::isfinite(light[0]) && std::isfinite(light[1]) && std::isfinite(light[2])))
        return fail(c, GPIS_INVALID_ARGUMENT, "light position must be finite");
    if (c->P.kind != GPIS_COV_SE)
        return fail(c, GPIS_UNSUPPORTED, "1D paths serve the stationary SE kernel");
    if (c->P.N != 8) return fail(c, GPIS_UNSUPPORTED, "1D flights are compiled for 8 impulses per cell");
    if (gpis_status s = mean_ready(c)) return s;
    if (n == 0) return GPIS_OK;
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(c, hipMemsetAsync(c->work, 0, sizeof(unsigned), st));
    unsigned threads = 256;
    int wpc = c->waves_per_cu > 0 ? c->waves_per_cu : 8;
    long long want = (long long)c->num_cus * wpc * 64 / threads;
    long long need = (n + threads - 1) / threads;
    unsigned blocks = (unsigned)(need < want ? need : (want < 1 ? 1 : want));
    Path1DIn pin = {seeds->px, seeds->py, seeds->spp, seeds->bounce, xxh1(seeds->scene_seed), max_bounces, phase,
                    bounce_seed, box};
    Path1DOut pout = {reply->t0,      reply->gx0,   reply->gy0,    reply->gz0,          reply->outcome0,
                      reply->flights, reply->hits,  reply->shadows, reply->visible,     reply->nee,
                      reply->t_last,  reply->outcome_last, reply->evals};
    launch_flight1d_kernel<true>(c, blocks, st, rays, (unsigned)n, ReplyOut{}, light, Nee1DOut{}, nullptr, nullptr, pin,
                                 pout);
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

// Launch the group kernel over the first *n_dev entries of the index list idx (n bounds them).
static gpis_status launch_flight1d_groups(gpis_ctx *c, const gpis_ray_soa *rays, int64_t n, const ReplyOut &out,
                                          const uint32_t *idx, const uint32_t *n_dev, hipStream_t st,
                                          unsigned *work = nullptr) {
    // the persistent grid's work counter: the context's, or one the caller cleared
    if (!work) {
        work = c->work;
        HIP_TRY(c, hipMemsetAsync(work, 0, sizeof(unsigned), st));
    }
    const unsigned threads = 256, groups_per_block = threads / kGroup;
    const long long want = (long long)c->num_cus * 2;  // two blocks of four waves per CU (2 waves per SIMD)
    const long long need = (n + groups_per_block - 1) / groups_per_block;
    const unsigned blocks = (unsigned)(need < want ? need : (want < 1 ? 1 : want));
#define GPIS_1DG_LAUNCH(PROC, FAST)                                                                              \
    hipLaunchKernelGGL((flight1d_group_kernel<8, PROC, FAST>), dim3(blocks), dim3(threads), 0, st, c->P_dev,          \
                       make_in(rays), rays->length_scale, out, work, c->counters, c->nf1d, idx, n_dev)
    const bool proc = c->P.mean_kind == GPIS_MEAN_PROCEDURAL;
    if (proc && c->fast1d) GPIS_1DG_LAUNCH(true, true);
    else if (proc) GPIS_1DG_LAUNCH(true, false);
    else if (c->fast1d) GPIS_1DG_LAUNCH(false, true);
    else GPIS_1DG_LAUNCH(false, false);
#undef GPIS_1DG_LAUNCH
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_trace_flights_1d(gpis_ctx *c, const gpis_ray_soa *rays, int64_t n, const gpis_reply_soa *reply,
                                  const float *light, const gpis_nee_out *nee, void *stream) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (!c->P.one_d) return fail(c, GPIS_INVALID_ARGUMENT, "the medium is not 1D_sampling");
    if (!rays) return fail(c, GPIS_INVALID_ARGUMENT, "null rays");
    if (n < 0 || n >= (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "ray count out of range");
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->near_t ||
        !rays->far_t || !rays->jitter || !rays->noise_seed || !rays->normal_seed || !rays->flags ||
        !rays->prev_value || !rays->prev_gx || !rays->prev_gy || !rays->prev_gz)
        return fail(c, GPIS_INVALID_ARGUMENT, "null ray field");
    if (!reply || !reply->t || !reply->value || !reply->gx || !reply->gy || !reply->gz || !reply->outcome)
        return fail(c, GPIS_INVALID_ARGUMENT, "null reply field");
    if (light && (!nee || !nee->gx || !nee->gy || !nee->gz || !nee->pdf))
        return fail(c, GPIS_INVALID_ARGUMENT, "NEE needs all four output arrays");
    if (c->P.N != 8) return fail(c, GPIS_UNSUPPORTED, "1D flights are compiled for 8 impulses per cell");
    if (gpis_status s = mean_ready(c)) return s;
    if (n == 0) return GPIS_OK;
    ReplyOut out = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals};
    Nee1DOut no = {};
    if (light) no = {nee->gx, nee->gy, nee->gz, nee->pdf};
    return launch_flight1d(c, rays, n, out, light, no, nullptr, nullptr, (hipStream_t)stream);
}

gpis_status gpis_get_counters(gpis_ctx *c, gpis_counters *o) {
    if (!c || !o) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    unsigned long long h[16];
    HIP_TRY(c, hipDeviceSynchronize());
    HIP_TRY(c, hipMemcpy(h, c->counters, sizeof h, hipMemcpyDeviceToHost));
    o->flights = h[0]; o->evals = h[1]; o->steps = h[2]; o->hits = h[3];
    o->exits = h[4]; o->fails = h[5]; o->invalid = h[6]; o->kernel_launches = h[7];
    o->spec_evals = h[8]; o->passes = h[9]; o->paired = h[10];
    return GPIS_OK;
}

gpis_status gpis_reset_counters(gpis_ctx *c) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    HIP_TRY(c, hipMemset(c->counters, 0, 16 * sizeof(unsigned long long)));
    return GPIS_OK;
}

gpis_status gpis_set_launch(gpis_ctx *c, int32_t waves_per_cu, int32_t block_size) {
    if (!c) return GPIS_INVALID_ARGUMENT;
    if (waves_per_cu < 0 || waves_per_cu > 32) return fail(c, GPIS_INVALID_ARGUMENT, "waves_per_cu out of range");
    if (block_size != 0 && (block_size < 64 || block_size > 256 || block_size % 64))
        return fail(c, GPIS_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
    c->waves_per_cu = waves_per_cu;
    c->block_size = block_size ? block_size : 256;
    return GPIS_OK;
}

}  // extern "C"

// ===================================================================== surrogate --
// Features of a conditioned 1D flight (order documented in include/gpis_surrogate.h).
__global__ void features1d_kernel(const FieldParams *__restrict__ Pg, RayIn in, const float *__restrict__ lscale, int n,
                                  float *__restrict__ feats) {
    const FieldParams &P = *Pg;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float ox = in.ox[r], oy = in.oy[r], oz = in.oz[r];
    const float dx = in.dx[r], dy = in.dy[r], dz = in.dz[r];
    const float near_t = in.near_t[r];
    float far_t = in.far_t[r];
    if (!isfinite(far_t)) far_t = near_t + 2000.f;
    float mgx, mgy, mgz;
    mean_grad_d(P, ox, oy, oz, mgx, mgy, mgz);
    float f[8];
    f[0] = in.pv[r];
    f[1] = dot3(in.pgx[r], in.pgy[r], in.pgz[r], dx, dy, dz);
    f[2] = lscale ? lscale[r] : P.length_scale;
    f[3] = P.amp_const * P.amp_const;
    f[4] = far_t - near_t;
    f[5] = (float)mean_value_d(P, ox, oy, oz);
    f[6] = dot3(mgx, mgy, mgz, dx, dy, dz);
    f[7] = P.kind == GPIS_COV_SE ? 1.0f : 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) feats[(size_t)r * 8 + i] = f[i];
}

// Training class of a plain-method flight: bin of s = (t - near) / (far - near), 31 for
// an exit, -1 for an invalid request.
__global__ void classes1d_kernel(RayIn in, const float *__restrict__ t, const uint32_t *__restrict__ outcome, int n,
                                 const mlp::Calibration *__restrict__ C, int32_t *__restrict__ cls) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t oc = outcome[r] & 0xFFu;
    int32_t c;
    if (oc == GPIS_INVALID) {
        c = -1;
    } else if (oc == GPIS_EXIT) {
        c = mlp::kBins;
    } else {
        const float near_t = in.near_t[r];
        float far_t = in.far_t[r];
        if (!isfinite(far_t)) far_t = near_t + 2000.f;
        float sv = (t[r] - near_t) / (far_t - near_t);
        sv = fminf(fmaxf(sv, 0.0f), 0.99999994f);
        // class 0 holds s <= 0 (an atom when edges[1] == 0: crossings at the origin);
        // otherwise the last bin whose lower edge is <= s
        c = 0;
        if (sv > 0.0f)
            for (int b = 1; b < mlp::kBins; ++b)
                if (C->edges[b] <= sv) c = b;
    }
    cls[r] = c;
}

__global__ void fill_u32_kernel(uint32_t *p, int n, uint32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct gpis_net {
    int device = 0;
    int num_cus = 0;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    float *part = nullptr;
    int64_t part_chunks = 0;
    mlp::bf16 *afwd = nullptr, *abwd = nullptr;
    mlp::Calibration *calib = nullptr;
    bool calibrated = false;
    mlp::bf16 *xt = nullptr, *dzt = nullptr;  // dzt and zt live inside the xt allocation
    float *zt = nullptr;
    int64_t ws_rows = 0;
    float *feats = nullptr;
    int64_t feats_rows = 0;
    uint32_t *fallback = nullptr, *n_fallback = nullptr;
    int64_t fb_rows = 0;
    float *loss = nullptr;
    int64_t step = 0;
    float *avg = nullptr;  // moving average of the parameters (gpis_net_set_averaging)
    float avg_beta = 0.0f;
    // served endpoints run on a side stream beside the fallback flights (fork and join
    // through two events; the side stream has the default priority)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
};

static gpis_status nfail(gpis_net *n, gpis_status s, const char *msg) {
    if (n) n->err = msg;
    return s;
}
#define NET_TRY(n, expr)                                                           \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            if (n) (n)->err = std::string(#expr) + ": " + hipGetErrorString(_e);   \
            return GPIS_DEVICE_ERROR;                                              \
        }                                                                          \
    } while (0)

static gpis_status net_pack(gpis_net *n, hipStream_t st) {
    const int total = mlp::kFwdElems + mlp::kBwdElems;
    hipLaunchKernelGGL(mlp::pack_kernel, dim3((total + 255) / 256), dim3(256), 0, st, n->theta, n->afwd, n->abwd);
    NET_TRY(n, hipGetLastError());
    return GPIS_OK;
}

template <typename T>
static hipError_t grow(T **p, int64_t *have, int64_t want, int64_t per_row) {
    if (*have >= want) return hipSuccess;
    if (*p) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        (void)hipFree(*p);
        *p = nullptr;
    }
    hipError_t e = hipMalloc((void **)p, (size_t)want * per_row * sizeof(T));
    if (e == hipSuccess) *have = want;
    else *have = 0;
    return e;
}

// Largest chunk of rows (<= 1024, a multiple of 64) dividing the batch.
static int chunk_rows_for(int64_t batch) {
    for (int c = 1024; c >= 64; c >>= 1)
        if (batch % c == 0) return c;
    return 0;
}

static gpis_status net_backprop(gpis_net *n, const float *feats, const int32_t *classes, int64_t batch,
                                float *loss_sum, hipStream_t st, int *chunks_out) {
    if (!n->calibrated) return nfail(n, GPIS_INVALID_ARGUMENT, "calibration not set");
    if (!feats || !classes) return nfail(n, GPIS_INVALID_ARGUMENT, "null features or classes");
    if (batch < 64 || batch % 64 != 0 || batch > (1LL << 24))
        return nfail(n, GPIS_INVALID_ARGUMENT, "batch must be a positive multiple of 64, at most 2^24");
    const int cr = chunk_rows_for(batch);
    const int chunks = (int)(batch / cr);
    // workspace per row: bf16 Xt and dZt, fp32 hidden pre-activations (counted in bf16 units)
    NET_TRY(n, grow(&n->xt, &n->ws_rows, batch, (int64_t)mlp::kXtRows + mlp::kDztRows + 2 * 4 * mlp::kWidth));
    n->dzt = n->xt + (size_t)mlp::kXtRows * batch;
    n->zt = (float *)(n->dzt + (size_t)mlp::kDztRows * batch);
    NET_TRY(n, grow(&n->part, &n->part_chunks, chunks, (int64_t)mlp::kParams));
    float *ls = loss_sum ? loss_sum : n->loss;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(32), 0, st, (uint32_t *)ls, 1, 0u);
    hipLaunchKernelGGL(mlp::train_fwd_bwd_kernel, dim3((unsigned)(batch / 64)), dim3(256), 0, st, feats, classes,
                       (int)batch, n->calib, n->afwd, n->abwd, n->theta, n->xt, n->zt, n->dzt, ls);
    NET_TRY(n, hipGetLastError());
    hipLaunchKernelGGL(mlp::wgrad_kernel, dim3((unsigned)chunks, mlp::kLayers), dim3(256), 0, st, n->xt, n->dzt,
                       (int)batch, cr, n->part);
    NET_TRY(n, hipGetLastError());
    *chunks_out = chunks;
    return GPIS_OK;
}

extern "C" {

gpis_status gpis_net_create(uint64_t seed, gpis_net **out) {
    if (!out) return GPIS_INVALID_ARGUMENT;
    *out = nullptr;
    gpis_net *n = new gpis_net();
    hipError_t e = hipGetDevice(&n->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n->num_cus, hipDeviceAttributeMultiprocessorCount, n->device);
    const size_t pb = sizeof(float) * mlp::kParams;
    if (e == hipSuccess) e = hipMalloc((void **)&n->theta, pb);
    if (e == hipSuccess) e = hipMalloc((void **)&n->m, pb);
    if (e == hipSuccess) e = hipMalloc((void **)&n->v, pb);
    if (e == hipSuccess) e = hipMalloc((void **)&n->grad, pb);
    if (e == hipSuccess) e = hipMalloc((void **)&n->afwd, sizeof(mlp::bf16) * mlp::kFwdElems);
    if (e == hipSuccess) e = hipMalloc((void **)&n->abwd, sizeof(mlp::bf16) * mlp::kBwdElems);
    if (e == hipSuccess) e = hipMalloc((void **)&n->calib, sizeof(mlp::Calibration));
    if (e == hipSuccess) e = hipMalloc((void **)&n->loss, 64);
    if (e == hipSuccess) e = hipMalloc((void **)&n->n_fallback, 64);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&n->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&n->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&n->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)mlp::infer_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)(mlp::kFwdElems * sizeof(mlp::bf16) +
                                                       (4 * mlp::kWidth + mlp::kClasses) * sizeof(float)));
    if (e != hipSuccess) {
        fprintf(stderr, "gpis_net_create: %s\n", hipGetErrorString(e));
        gpis_net_destroy(n);
        return GPIS_DEVICE_ERROR;
    }
    // He-uniform weights, zero biases
    std::vector<float> theta(mlp::kParams, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < mlp::kLayers; ++l) {
        const float lim = sqrtf(6.0f / (float)mlp::layer_in(l));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < mlp::layer_in(l) * mlp::layer_out(l); ++i) theta[mlp::w_off(l) + i] = U(rng);
    }
    gpis_status s = gpis_net_set_params(n, theta.data(), mlp::kParams);
    if (s != GPIS_OK) {
        gpis_net_destroy(n);
        return s;
    }
    *out = n;
    return GPIS_OK;
}

gpis_status gpis_net_destroy(gpis_net *n) {
    if (!n) return GPIS_INVALID_ARGUMENT;
    (void)hipDeviceSynchronize();
    void *ptrs[] = {n->theta, n->m, n->v, n->grad, n->part, n->afwd, n->abwd, n->calib,
                    n->xt, n->feats, n->fallback, n->n_fallback, n->loss, n->avg};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (n->fork) (void)hipEventDestroy(n->fork);
    if (n->join) (void)hipEventDestroy(n->join);
    if (n->side) (void)hipStreamDestroy(n->side);
    delete n;
    return GPIS_OK;
}

const char *gpis_net_last_error(const gpis_net *n) { return n ? n->err.c_str() : "null network"; }

gpis_status gpis_net_set_averaging(gpis_net *n, float beta) {
    if (!n) return GPIS_INVALID_ARGUMENT;
    if (!(beta >= 0.0f && beta < 1.0f)) return nfail(n, GPIS_INVALID_ARGUMENT, "beta must lie in [0, 1)");
    if (beta > 0.0f) {
        const size_t pb = sizeof(float) * mlp::kParams;
        if (!n->avg) NET_TRY(n, hipMalloc((void **)&n->avg, pb));
        NET_TRY(n, hipMemcpy(n->avg, n->theta, pb, hipMemcpyDeviceToDevice));  // the average starts here
    }
    n->avg_beta = beta;
    return GPIS_OK;
}

gpis_status gpis_net_use_average(gpis_net *n) {
    if (!n) return GPIS_INVALID_ARGUMENT;
    if (!n->avg || !(n->avg_beta > 0.0f)) return nfail(n, GPIS_INVALID_ARGUMENT, "averaging is not enabled");
    NET_TRY(n, hipMemcpy(n->theta, n->avg, sizeof(float) * mlp::kParams, hipMemcpyDeviceToDevice));
    n->avg_beta = 0.0f;
    gpis_status s = net_pack(n, 0);
    if (s != GPIS_OK) return s;
    NET_TRY(n, hipDeviceSynchronize());
    return GPIS_OK;
}

gpis_status gpis_net_set_params(gpis_net *n, const float *params, int64_t count) {
    if (!n || !params) return nfail(n, GPIS_INVALID_ARGUMENT, "null argument");
    if (count != mlp::kParams) return nfail(n, GPIS_INVALID_ARGUMENT, "parameter count must be 54816");
    const size_t pb = sizeof(float) * mlp::kParams;
    NET_TRY(n, hipMemcpy(n->theta, params, pb, hipMemcpyHostToDevice));
    NET_TRY(n, hipMemset(n->m, 0, pb));
    NET_TRY(n, hipMemset(n->v, 0, pb));
    n->step = 0;
    n->avg_beta = 0.0f;  // new parameters end any averaging
    gpis_status s = net_pack(n, 0);
    if (s != GPIS_OK) return s;
    NET_TRY(n, hipDeviceSynchronize());
    return GPIS_OK;
}

gpis_status gpis_net_get_params(gpis_net *n, float *params, int64_t count) {
    if (!n || !params) return nfail(n, GPIS_INVALID_ARGUMENT, "null argument");
    if (count != mlp::kParams) return nfail(n, GPIS_INVALID_ARGUMENT, "parameter count must be 54816");
    NET_TRY(n, hipDeviceSynchronize());
    NET_TRY(n, hipMemcpy(params, n->theta, sizeof(float) * mlp::kParams, hipMemcpyDeviceToHost));
    return GPIS_OK;
}

gpis_status gpis_net_set_calibration(gpis_net *n, const gpis_net_calibration *cal) {
    if (!n || !cal) return nfail(n, GPIS_INVALID_ARGUMENT, "null argument");
    mlp::Calibration C;
    for (int i = 0; i < mlp::kIn; ++i) {
        C.mean[i] = cal->feat_mean[i];
        C.inv_std[i] = cal->feat_inv_std[i];
        C.lo[i] = cal->hull_lo[i];
        C.hi[i] = cal->hull_hi[i];
        if (!std::isfinite(C.mean[i]) || !std::isfinite(C.inv_std[i]) || !(C.lo[i] <= C.hi[i]))
            return nfail(n, GPIS_INVALID_ARGUMENT, "non-finite standardisation or empty hull");
    }
    if (cal->edges[0] != 0.0f || cal->edges[mlp::kBins] != 1.0f)
        return nfail(n, GPIS_INVALID_ARGUMENT, "bin edges must run from 0 to 1");
    for (int b = 0; b <= mlp::kBins; ++b) {
        C.edges[b] = cal->edges[b];
        // strictly increasing, except that edges[1] == 0 makes class 0 an atom at s = 0
        if (b > 0 && !(C.edges[b] > C.edges[b - 1]) && !(b == 1 && C.edges[1] == 0.0f))
            return nfail(n, GPIS_INVALID_ARGUMENT, "bin edges must increase");
    }
    C.entropy_max = cal->entropy_max;
    NET_TRY(n, hipMemcpy(n->calib, &C, sizeof C, hipMemcpyHostToDevice));
    n->calibrated = true;
    return GPIS_OK;
}

gpis_status gpis_net_features_1d(gpis_ctx *c, const gpis_ray_soa *rays, int64_t n, float *feats, void *stream) {
    if (!c || !rays || !feats) return fail(c, GPIS_INVALID_ARGUMENT, "null argument");
    if (!c->P.one_d) return fail(c, GPIS_INVALID_ARGUMENT, "the medium is not 1D_sampling");
    if (n < 0 || n >= (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "ray count out of range");
    if (gpis_status s = mean_ready(c)) return s;
    if (n == 0) return GPIS_OK;
    hipLaunchKernelGGL(features1d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       c->P_dev, make_in(rays), rays->length_scale, (int)n, feats);
    HIP_TRY(c, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_net_classes_1d(gpis_net *net, const gpis_ray_soa *rays, const gpis_reply_soa *reply, int64_t n,
                                int32_t *classes, void *stream) {
    if (!net || !rays || !reply || !reply->t || !reply->outcome || !classes)
        return nfail(net, GPIS_INVALID_ARGUMENT, "null argument");
    if (!net->calibrated) return nfail(net, GPIS_INVALID_ARGUMENT, "calibration not set");
    if (n < 0 || n >= (1LL << 31)) return nfail(net, GPIS_INVALID_ARGUMENT, "ray count out of range");
    if (n == 0) return GPIS_OK;
    hipLaunchKernelGGL(classes1d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       make_in(rays), reply->t, reply->outcome, (int)n, net->calib, classes);
    NET_TRY(net, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_net_gradient(gpis_net *n, const float *feats, const int32_t *classes, int64_t batch, float *grad,
                              float *loss_sum, void *stream) {
    if (!n || !grad) return nfail(n, GPIS_INVALID_ARGUMENT, "null argument");
    hipStream_t st = (hipStream_t)stream;
    int chunks = 0;
    gpis_status s = net_backprop(n, feats, classes, batch, loss_sum, st, &chunks);
    if (s != GPIS_OK) return s;
    hipLaunchKernelGGL(mlp::adam_kernel, dim3((mlp::kParams + 255) / 256), dim3(256), 0, st, n->part, chunks,
                       1.0f / (float)batch, n->theta, n->m, n->v, grad, 0.0f, 1.0f, 1.0f, 0, (float *)nullptr,
                       0.0f);
    NET_TRY(n, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_net_train_step(gpis_net *n, const float *feats, const int32_t *classes, int64_t batch, float lr,
                                float *loss_sum, void *stream) {
    if (!n) return GPIS_INVALID_ARGUMENT;
    if (!(lr > 0.0f) || !std::isfinite(lr)) return nfail(n, GPIS_INVALID_ARGUMENT, "learning rate must be positive");
    hipStream_t st = (hipStream_t)stream;
    int chunks = 0;
    gpis_status s = net_backprop(n, feats, classes, batch, loss_sum, st, &chunks);
    if (s != GPIS_OK) return s;
    n->step += 1;
    const float bc1 = (float)(1.0 / (1.0 - pow(0.9, (double)n->step)));
    const float bc2 = (float)(1.0 / (1.0 - pow(0.999, (double)n->step)));
    hipLaunchKernelGGL(mlp::adam_kernel, dim3((mlp::kParams + 255) / 256), dim3(256), 0, st, n->part, chunks,
                       1.0f / (float)batch, n->theta, n->m, n->v, n->grad, lr, bc1, bc2, 1,
                       n->avg_beta > 0.0f ? n->avg : (float *)nullptr, n->avg_beta);
    NET_TRY(n, hipGetLastError());
    return net_pack(n, st);
}

static gpis_status net_infer(gpis_net *n, const float *feats, int64_t rows, const float *near_t, const float *far_t,
                             const float *jitter, const uint32_t *flags, const mlp::InferOut &o, hipStream_t st) {
    const size_t smem = mlp::kFwdElems * sizeof(mlp::bf16) + (4 * mlp::kWidth + mlp::kClasses) * sizeof(float);
    const int64_t groups = (rows + 16 * mlp::kNT - 1) / (16 * mlp::kNT);
    int64_t blocks = (groups + mlp::kInferWaves - 1) / mlp::kInferWaves;
    if (blocks > n->num_cus) blocks = n->num_cus;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(mlp::infer_kernel, dim3((unsigned)blocks), dim3(mlp::kInferWaves * 64), smem, st, feats,
                       near_t, far_t, jitter, flags, (int)rows, n->calib, n->afwd, n->theta, o);
    NET_TRY(n, hipGetLastError());
    return GPIS_OK;
}

gpis_status gpis_net_forward(gpis_net *n, const float *feats, int64_t rows, float *probs, void *stream) {
    if (!n || !feats || !probs) return nfail(n, GPIS_INVALID_ARGUMENT, "null argument");
    if (!n->calibrated) return nfail(n, GPIS_INVALID_ARGUMENT, "calibration not set");
    if (rows < 0 || rows >= (1LL << 31)) return nfail(n, GPIS_INVALID_ARGUMENT, "row count out of range");
    if (rows == 0) return GPIS_OK;
    mlp::InferOut o = {};
    o.probs = probs;
    return net_infer(n, feats, rows, nullptr, nullptr, nullptr, nullptr, o, (hipStream_t)stream);
}

}  // extern "C"

// The media the learned sampler serves: conditioned (Renewal, Renewal+) single-level
// stationary SE flights with a non-zero amplitude, the law the served endpoint assumes.
static gpis_status surrogate_medium_ok(gpis_ctx *c) {
    if (!c->P.one_d) return fail(c, GPIS_INVALID_ARGUMENT, "the medium is not 1D_sampling");
    if (c->P.ctxt != GPIS_CTX_RENEWAL && c->P.ctxt != GPIS_CTX_RENEWAL_PLUS)
        return fail(c, GPIS_UNSUPPORTED, "the learned sampler serves Renewal and Renewal+ flights");
    if (c->P.kind != GPIS_COV_SE || c->P.multires)
        return fail(c, GPIS_UNSUPPORTED, "the learned sampler serves the stationary single-level SE kernel");
    if (!(c->P.amp_const > 0.f)) return fail(c, GPIS_UNSUPPORTED, "the learned sampler needs sigma > 0");
    return GPIS_OK;
}

// After the inference kernel on `st`: the fallback flights (the plain method, a group of
// lanes per flight) on `st`, and beside them the served rows' endpoints on the network's
// side stream (fork and join through two events; the side stream has the default priority).
static gpis_status launch_served_and_fallback(gpis_ctx *c, gpis_net *net, const gpis_ray_soa *rays, int64_t n,
                                              const ReplyOut &out, const uint32_t *fallback, const uint32_t *n_fb,
                                              hipStream_t st, unsigned *work) {
    HIP_TRY(c, hipEventRecord(net->fork, st));
    gpis_status s = launch_flight1d_groups(c, rays, n, out, fallback, n_fb, st, work);
    if (s != GPIS_OK) return s;
    HIP_TRY(c, hipStreamWaitEvent(net->side, net->fork, 0));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (c->P.mean_kind == GPIS_MEAN_PROCEDURAL)
        hipLaunchKernelGGL(served_finish_kernel<true>, dim3(blocks), dim3(256), 0, net->side, c->P_dev, make_in(rays),
                           rays->length_scale, (unsigned)n, out);
    else
        hipLaunchKernelGGL(served_finish_kernel<false>, dim3(blocks), dim3(256), 0, net->side, c->P_dev,
                           make_in(rays), rays->length_scale, (unsigned)n, out);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(net->join, net->side));
    HIP_TRY(c, hipStreamWaitEvent(st, net->join, 0));
    return GPIS_OK;
}

extern "C" {

gpis_status gpis_surrogate_flights_1d(gpis_ctx *c, gpis_net *net, const gpis_ray_soa *rays, int64_t n,
                                      const gpis_reply_soa *reply, uint32_t *n_fallback, void *stream) {
    if (!c || !net) return GPIS_INVALID_ARGUMENT;
    if (!net->calibrated) return nfail(net, GPIS_INVALID_ARGUMENT, "calibration not set");
    if (gpis_status ms = surrogate_medium_ok(c)) return ms;
    if (!rays || !reply || !reply->t || !reply->value || !reply->gx || !reply->gy || !reply->gz || !reply->outcome)
        return fail(c, GPIS_INVALID_ARGUMENT, "null ray or reply field");
    if (n < 0 || n >= (1LL << 31)) return fail(c, GPIS_INVALID_ARGUMENT, "ray count out of range");
    if (gpis_status ms = mean_ready(c)) return ms;
    if (n_fallback) *n_fallback = 0;
    if (n == 0) return GPIS_OK;
    if (c->P.N != 8) return fail(c, GPIS_UNSUPPORTED, "1D flights are compiled for 8 impulses per cell");
    hipStream_t st = (hipStream_t)stream;
    NET_TRY(net, grow(&net->feats, &net->feats_rows, n, (int64_t)mlp::kIn));
    NET_TRY(net, grow(&net->fallback, &net->fb_rows, n, (int64_t)1));
    gpis_status s = gpis_net_features_1d(c, rays, n, net->feats, stream);
    if (s != GPIS_OK) return s;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(32), 0, st, net->n_fallback, 1, 0u);
    mlp::InferOut o = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals,
                       nullptr, net->fallback, net->n_fallback};
    s = net_infer(net, net->feats, n, (const float *)rays->near_t, (const float *)rays->far_t,
                  (const float *)rays->jitter, (const uint32_t *)rays->flags, o, st);
    if (s != GPIS_OK) return s;
    ReplyOut out = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals};
    // the fallback rays are few and their flights long: a group of lanes per flight
    s = launch_served_and_fallback(c, net, rays, n, out, net->fallback, net->n_fallback, st, nullptr);
    if (s != GPIS_OK) return s;
    if (n_fallback) {
        HIP_TRY(c, hipMemcpyAsync(n_fallback, net->n_fallback, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
    }
    return GPIS_OK;
}

}  // extern "C"

// ------------------------------------------------ allocation-free surrogate flights --
// Scratch of gpis_surrogate_flights_1d_noalloc: the features of the rays, the list of
// rays that fall back to the plain method, its length and the plain kernel's work
// counter.  One copy per GPU, held until the process ends; host code uses the address
// hipGetSymbolAddress returns on each device (the array's name is only a host stand-in).
constexpr int64_t kNoAllocMaxRays = 1 << 18;
constexpr size_t kNoAllocFeatBytes = (size_t)kNoAllocMaxRays * mlp::kIn * sizeof(float);  // 8 MiB
constexpr size_t kNoAllocListBytes = (size_t)kNoAllocMaxRays * sizeof(uint32_t);           // 1 MiB
constexpr size_t kNoAllocBytes = kNoAllocFeatBytes + kNoAllocListBytes + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static unsigned char *g_scratch[kNoAllocMaxDevices];
static std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device (looked up once per device).
static hipError_t noalloc_scratch(unsigned char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

extern "C" {

gpis_status gpis_surrogate_flights_1d_noalloc(gpis_ctx *c, gpis_net *net, const gpis_ray_soa *rays, int64_t n,
                                              const gpis_reply_soa *reply, uint32_t *n_fallback, void *stream) {
    if (!c || !net) return GPIS_INVALID_ARGUMENT;
    if (!net->calibrated) return nfail(net, GPIS_INVALID_ARGUMENT, "calibration not set");
    if (gpis_status ms = surrogate_medium_ok(c)) return ms;
    if (!rays || !reply || !reply->t || !reply->value || !reply->gx || !reply->gy || !reply->gz || !reply->outcome)
        return fail(c, GPIS_INVALID_ARGUMENT, "null ray or reply field");
    if (n < 0 || n > kNoAllocMaxRays)
        return fail(c, GPIS_INVALID_ARGUMENT, "the allocation-free entry point takes at most 262144 rays per call");
    if (gpis_status ms = mean_ready(c)) return ms;
    if (n_fallback) *n_fallback = 0;
    if (n == 0) return GPIS_OK;
    if (c->P.N != 8) return fail(c, GPIS_UNSUPPORTED, "1D flights are compiled for 8 impulses per cell");
    int dev = -1;
    HIP_TRY(c, hipGetDevice(&dev));
    if (dev != c->device || dev != net->device)
        return fail(c, GPIS_INVALID_ARGUMENT, "the context, the network and the current device must agree");
    unsigned char *sp = nullptr;
    HIP_TRY(c, noalloc_scratch(&sp));
    float *feats = (float *)sp;
    uint32_t *fallback = (uint32_t *)(sp + kNoAllocFeatBytes);
    uint32_t *counts = (uint32_t *)(sp + kNoAllocFeatBytes + kNoAllocListBytes);  // [0] fallback rays, [1] work
    hipStream_t st = (hipStream_t)stream;
    gpis_status s = gpis_net_features_1d(c, rays, n, feats, stream);
    if (s != GPIS_OK) return s;
    // the counters start at zero (a fill kernel: the entry point records no memset)
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(32), 0, st, counts, 2, 0u);
    HIP_TRY(c, hipGetLastError());
    mlp::InferOut o = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals,
                       nullptr, fallback, counts};
    s = net_infer(net, feats, n, (const float *)rays->near_t, (const float *)rays->far_t,
                  (const float *)rays->jitter, (const uint32_t *)rays->flags, o, st);
    if (s != GPIS_OK) return s;
    ReplyOut out = {reply->t, reply->value, reply->gx, reply->gy, reply->gz, reply->outcome, reply->evals};
    s = launch_served_and_fallback(c, net, rays, n, out, fallback, counts, st, counts + 1);
    if (s != GPIS_OK) return s;
    if (n_fallback) {
        HIP_TRY(c, hipMemcpyAsync(n_fallback, counts, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
    }
    return GPIS_OK;
}

}  // extern "C"
