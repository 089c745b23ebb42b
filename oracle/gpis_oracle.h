/*
 * gpis_oracle.h — CPU restatement of the sparse-convolution GPIS ray-flight path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it.  The product (gpis_extend_amd, libgpis_hip.so) never calls it.
 *
 * It restates, function by function, the reference's CPU algorithm with the
 * same float/double choices:
 *   src/core/math/MathUtil.hpp:179-228            xxhash32, Bernoulli
 *   src/core/sampling/UniformSampler.hpp:17-75     PCG32 (XSH-RR, inc 1)
 *   src/core/math/BitManip.hpp:47-50               normalizedUint
 *   src/core/math/TangentFrame.hpp:74-85           Duff et al. basis
 *   src/core/sampling/Gaussian.cpp:22-34,89-104    Box-Muller
 *   src/core/math/GPFunctions.cpp:555-875          SE splat / covariance kernels
 *   src/core/math/GPFunctions.cpp:1225-1345,1600-1760  nonstationary covariance
 *   src/core/math/GPFunctions.hpp:952-996          spherical mean
 *   src/core/math/GPFunctions.cpp:250-298, ProceduralMeanSdf.cpp:11-38,
 *   src/core/math/SdfFunctionsQueries.cpp:14-466   procedural SDF means
 *   src/core/math/SparseConvolutionNoise.cpp       realization (all)
 *   src/core/media/SparseConvolutionNoiseMedium.cpp:199-347  prepare/finish
 *   src/core/media/GaussianProcessMedium.cpp:222-440         outcome rules
 */
#ifndef GPIS_ORACLE_H_
#define GPIS_ORACLE_H_

#include <stdint.h>

/* A binary32 float, also in the FP64 truth build of the oracle (gpis_oracle.c compiled
 * with ORC_FP64_TRUTH, where the word float names double): the layout of every struct
 * below is the same in both builds. */
typedef __typeof__(1.0f) orc_f32;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- leaves ---------------------------------------------------------- */
uint32_t orc_xxhash32_1(uint32_t p);
uint32_t orc_xxhash32_2(uint32_t x, uint32_t y);
uint32_t orc_xxhash32_4(uint32_t x, uint32_t y, uint32_t z, uint32_t w);
/* Impulses of one 3D cell: out[4*k..4*k+3] = (x, y, z, weight).
 * (i, j, k) are the cell's x, y, z indices cast to uint32. */
void orc_cell3d_impulses(uint32_t i, uint32_t j, uint32_t k, uint32_t seed,
                         uint32_t n, float *out);
/* Impulses of one 1D cell: out[2*k..2*k+1] = (t, weight). */
void orc_cell1d_impulses(uint32_t i, uint32_t seed, uint32_t n, float *out);
/* Box-Muller pair after UniformSampler::set_state(state). out[2*m], m < npairs. */
void orc_standard_normals(uint64_t state, int npairs, double *out);
void orc_tangent_frame(const float n[3], float out9[9]); /* tangent, bitangent, normal */

/* ---- medium description --------------------------------------------- */
enum { ORC_COV_SE = 0, ORC_COV_PROC_NS = 1 };
enum { ORC_MEAN_SPHERE = 0, ORC_MEAN_BOX = 1, ORC_MEAN_PLANE = 2, ORC_MEAN_GRID = 3, ORC_MEAN_PROC = 4 };
/* SdfFunctions::Function (SdfFunctions.hpp:13-22) */
enum { ORC_SDF_KNOB = 0, ORC_SDF_KNOB_INNER = 1, ORC_SDF_KNOB_OUTER = 2, ORC_SDF_TWO_SPHERES = 3,
       ORC_SDF_PLANE = 4, ORC_SDF_ROUNDED_U_GROOVE = 5, ORC_SDF_TWO_SLABS = 6, ORC_SDF_THREE_SLABS = 7 };
enum { ORC_CTX_GLOBAL = 0, ORC_CTX_RENEWAL = 1, ORC_CTX_RENEWAL_PLUS = 2, ORC_CTX_NONE = 3 };
enum { ORC_SCHEME_UNI = 0, ORC_SCHEME_NEE = 1, ORC_SCHEME_MIS = 2 };
enum { ORC_LS_CONST = 0, ORC_LS_SINCOS = 1, ORC_LS_NOISE = 2 };
enum { ORC_VAR_CONST = 0, ORC_VAR_NOISE = 1 };
/* ProceduralNoise::NoiseType (GPFunctions.hpp:655-662) */
enum { ORC_NOISE_BOTTOM_TOP = 0, ORC_NOISE_LEFT_RIGHT = 1, ORC_NOISE_FRONT_BACK = 2,
       ORC_NOISE_BOTTOM_TOP_LEFT_RIGHT = 3, ORC_NOISE_SANDSTONE = 4, ORC_NOISE_RUST = 5 };

typedef struct orc_medium_desc {
    /* covariance (SquaredExponentialCovariance JSON: sigma, lengthScale, localScale, aniso) */
    int32_t cov_kind;
    orc_f32 sigma, length_scale, kernel_scale;
    orc_f32 aniso[3];
    /* ProceduralNonstationaryCovariance: ls field, multiResolutionGrid */
    int32_t ns_multires;
    int32_t ns_ls_field;
    orc_f32 ns_ls_params[4];   /* SINCOS: base, amplitude, frequency (x0, x2) */
    orc_f32 ns_ls_max;         /* ProceduralVector::maxVal() of the ls field */
    orc_f32 ns_var;            /* constant variance field */
    /* mean */
    int32_t mean_kind;
    orc_f32 mean_radius;
    double mean_center[3];
    double mean_half[3];     /* BOX half extents */
    double plane_normal[3];  /* PLANE: mean = dot(n, p) - offset */
    double plane_offset;
    /* SparseConvolutionNoiseMedium JSON */
    uint32_t global_seed;
    orc_f32 impulse_density;
    int32_t single_realization;
    int32_t ctxt;
    int32_t one_d;
    int32_t scheme;
    int32_t correlation_xy;
    orc_f32 step_size;
    uint32_t min_step;
    int32_t max_bounce;
    /* TabulatedMean over a dense voxel grid (the HostMeanGrid fields of
     * src/gpu/cuda/RaySegmentBatch.hpp:19-30); the voxels stay owned by the caller. */
    const orc_f32 *grid;          /* x fastest */
    int32_t grid_origin[3], grid_dims[3];
    orc_f32 grid_w2i[12];         /* row-major 3x4 world -> index */
    orc_f32 grid_clamp_min[3], grid_clamp_max[3];
    orc_f32 grid_offset, grid_scale;
    int32_t grid_volume;
    /* "isotropic_3D_sampling": 3D noise in the isotropic ray space of each flight
     * (evaluateNoise3DIsotropicRayNormalized, SparseConvolutionNoise.cpp:345-377) */
    int32_t iso3d;
    /* ProceduralMean over an SdfFunctions function (GPFunctions.cpp:250-298): "func",
     * rows 0-2 of _invConfigTransform (row-major), "min", "scale", "offset" */
    int32_t sdf_function;
    orc_f32 sdf_inverse_transform[12];
    orc_f32 sdf_min, sdf_scale, sdf_offset;
    /* ProceduralNonstationaryCovariance fields of type "noise" (ProceduralNoise "var",
     * ProceduralNoiseVec "ls"; GPFunctions.hpp:647-830): ns_var_field ORC_VAR_NOISE and/or
     * ns_ls_field ORC_LS_NOISE; the noise type and its JSON min, max, start, end, min2, max2,
     * start2, end2 */
    int32_t ns_var_field;
    int32_t ns_var_noise, ns_ls_noise;
    double ns_var_noise_params[8], ns_ls_noise_params[8];
    /* "useAnisoMtx" and "anisoMtx" (row-major) of the SE covariance (GPFunctions.cpp:663-681) */
    int32_t use_aniso_mtx;
    orc_f32 aniso_mtx[9];
} orc_medium_desc;

/* Dense-grid density at a world point (trilinear, clamped), and TabulatedMean::mean
 * and dmean_da at double points (GPFunctions.cpp:165-190). */
float orc_grid_density(const orc_medium_desc *d, const float p[3]);
double orc_grid_mean(const orc_medium_desc *d, const double a[3]);
void orc_grid_mean_grad(const orc_medium_desc *d, const double a[3], double g[3]);

/* SdfFunctions::evaluateJet at a local point (SdfFunctionsQueries.cpp:444-450): value,
 * gradient into g.  Only for the functions with queries (knob, two_spheres, two_slabs,
 * three_slabs, rounded_u_groove). */
double orc_sdf_jet(int fn, const double p[3], double g[3]);
/* SdfFunctions::eval, the float API (SdfFunctions.hpp:64-75, SdfFunctionsQueries.cpp:374-394). */
float orc_sdf_eval(int fn, const float p[3]);
/* ProceduralMean::mean and dmean_da (GPFunctions.cpp:282-298, ProceduralMeanSdf.cpp:32-38)
 * at a double point; g may be NULL. */
double orc_sdf_mean(const orc_medium_desc *d, const double a[3], double g[3]);

/* One flight request: the fields of GpRaySegmentRequest plus MediumState. */
typedef struct orc_ray {
    orc_f32 origin[3], dir[3];
    orc_f32 near_t, far_t, jitter;
    uint32_t pss[4];          /* RayInfo::pixelSampleSegment */
    uint32_t scene_seed;
    int32_t first_scatter;
    orc_f32 last_val;           /* MediumState::lastVal */
    double last_aniso[3];     /* MediumState::lastAniso */
    int32_t bounce;
} orc_ray;

enum { ORC_EXIT = 0, ORC_HIT = 1, ORC_FAIL = 2 };
enum { ORC_FAIL_NONE = 0, ORC_FAIL_OUTWARD = 1, ORC_FAIL_ZERO_GRAD = 2,
       ORC_FAIL_NONFINITE = 3, ORC_FAIL_BOUNCE_LIMIT = 4 };

typedef struct orc_flight {
    double t;            /* sampled distance (hit) or farT (exit) */
    orc_f32 value;         /* MediumState::lastVal after the flight */
    orc_f32 grad[3];       /* sample.aniso: gradient at the hit, or at farT on exit */
    int32_t outcome;     /* ORC_EXIT / ORC_HIT / ORC_FAIL */
    int32_t fail_reason;
    int32_t steps;       /* lattice steps taken */
    int32_t evals;       /* field evaluations performed (value or value+gradient) */
    int32_t scheme;      /* sparseConv1DSamplingScheme at the endpoint */
    orc_f32 min_abs_f;     /* smallest |f| of any evaluation whose sign decided the march */
    orc_f32 bracket[2];    /* f at the lattice points bracketing the crossing (pf, fc) */
} orc_flight;

/* ---- field API ------------------------------------------------------- */
typedef struct orc_field orc_field;
orc_field *orc_field_create(const orc_medium_desc *desc);
void orc_field_destroy(orc_field *f);
/* Realization seed for a RayInfo (computeSeed). */
uint32_t orc_compute_seed(const orc_field *f, const uint32_t pss[4], uint32_t scene_seed);
/* Unconditioned normalised noise + world gradient (evaluateNoise3D), per point. */
void orc_noise3d(const orc_field *f, const float *p, uint32_t seed, float out4[4]);
/* f(p) and grad f(p) as evaluateValue/evaluateGradient with no conditioning. */
void orc_eval_points_dir(const orc_field *f, const float *xyz, int64_t n, uint32_t seed, const float dir[3],
                         float *out4, int nthreads);
void orc_eval_points(const orc_field *f, const float *xyz, int64_t n,
                     uint32_t seed, float *out4, int nthreads);
/* World-space cell size used by the 3D grid (splattingKernelRadius(false, 1)). */
float orc_kernel_radius(const orc_field *f);
/* One full flight (prepare + finish + resumeDistance outcome rules). */
void orc_trace_flight(const orc_field *f, const orc_ray *ray, orc_flight *out);
void orc_trace_flights(const orc_field *f, const orc_ray *rays, int64_t n,
                       orc_flight *out, int nthreads);
/* Test hook: the same flight with +-eps added to every marched value (hashed signs
 * from pseed).  A result that changes under eps of the size of FP32 evaluation
 * error is ill-conditioned (grazing); tests use it to classify disagreements. */
void orc_trace_flight_perturbed(const orc_field *f, const orc_ray *ray, float eps, uint32_t pseed,
                                orc_flight *out);
/* Conditioning check: evaluate value and gradient of the conditioned realization
 * at q, after conditioning at ray origin on (last_val, last_aniso). */
void orc_conditioned_eval(const orc_field *f, const orc_ray *ray, const float q[3],
                          float t_segment, float out4[4]);
/* 1D NEE (neeGrad / neePDF) after conditioning on the ray's origin state. */
int orc_transverse_moments(const orc_field *f, const orc_ray *ray, const float p[3], double mean[2], double *var,
                           int *scheme);
void orc_nee(const orc_field *f, const orc_ray *ray, const float normal[3],
             const float p[3], float t_segment, float grad_out[3], float *pdf_out);

/* ---- 1D paths: flights chained through the hit-time phase ------------ */
/* A path of a 1D_sampling medium inside the box [-box, box]^3 lit by one point light
 * (Point::sampleDirect, src/core/primitives/Point.cpp:98-106, emission 1), following
 * PathTracer's medium branch (PathTracer.cpp:380-438): flight b of the path has
 * pixelSampleSegment.w = w0 + b (w0 = pss[3] of the path's first ray); at a hit the
 * phase is BRDFPhaseFunction (BRDFPhaseFunction.cpp:29-100) with a mirror
 * (MirrorBsdf.cpp:30-113) or Lambert (LambertBsdf.cpp:27-47) BSDF of albedo 1;
 * next-event estimation is volumeLightSample (TraceBase.cpp:798-839): f = phase.eval
 * (the mirror's neePDF of the half-vector normal, zero under the UNI scheme), a
 * shadow flight conditioned on (0, neeGrad) (mirror) or (0, hit gradient) (Lambert)
 * toward the light over min(distance, box exit) whose any crossing occludes
 * (resumeTransmittance, GaussianProcessMedium.cpp:394-421), and f / |light - p|^2 is
 * added when unoccluded.  The next flight starts at the hit conditioned on (0,
 * gradient) along the phase-sampled direction to the box exit.  The path sampler's
 * draws are replaced by PCG32 streams seeded by xxhash32(path, b, bounce_seed, s) + 1:
 * stream 0 of flight b >= 1 gives (u1, u2, jitter) of its phase sample and flight,
 * stream 1 of flight b the jitter of its shadow flight. */
enum { ORC_PHASE_MIRROR = 0, ORC_PHASE_LAMBERT = 1 };
typedef struct orc_path {
    orc_f32 t0, grad0[3];       /* first flight */
    int32_t outcome0, fail0;
    int32_t flights, hits;    /* path flights, hits among them */
    int32_t shadows, visible; /* shadow flights, unoccluded among them */
    orc_f32 nee;                /* sum over scattering events of f * V / |light - p|^2 */
    orc_f32 t_last;             /* last path flight */
    int32_t outcome_last, fail_last;
    int32_t evals;            /* field evaluations, shadow flights included */
    orc_f32 pos_last[3], dir_last[3]; /* origin and direction of the last path flight */
} orc_path;
/* Up to 1 + max_bounces flights per path (PathTracer's maxBounces = 1 + max_bounces);
 * ls: optional per-path SE length scale; light: NULL for no next-event estimation. */
void orc_trace_paths1d(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, int64_t n,
                       int32_t max_bounces, int32_t phase, uint32_t bounce_seed, float box,
                       const float *light, orc_path *out, int nthreads);

/* ---- 3D paths: flights chained through diffuse bounces ---------------- */
/* The device path kernel's chaining (gpis_trace_paths): every hit continues along a
 * cosine-weighted diffuse bounce about the unit gradient (PCG32 seeded by
 * xxhash32(path, bounce, bounce_seed, 0) + 1 gives u1, u2 and the jitter) from the hit to
 * the exit of [-box, box]^3, conditioned on (0, the hit's gradient); at most
 * 1 + max_bounces flights.  seeded: flight b uses pixelSampleSegment.w = w0 + b.
 * perturb (test hook, 0 = off): +-perturb with hashed signs (from pseed, the path and the
 * flight) added to every marched value, as orc_trace_flight_perturbed. */
void orc_trace_paths3d(const orc_field *f, const orc_ray *rays, int64_t n, int32_t max_bounces,
                       uint32_t bounce_seed, float box, int32_t seeded, float perturb, uint32_t pseed,
                       orc_path *out, int nthreads);
void orc_sincos_turn(float u, float *s, float *c);
void orc_diffuse_direction(uint32_t ray, uint32_t bounce, uint32_t bseed, const float n[3], float d[3],
                           float *jit);

/* Per-ray SE length scale (1D media whose rays each draw their own l): the same
 * flights and NEE as above, each ray on the field of `desc` with length_scale =
 * ls[i].  `normals`/`points`/`t_segment` are per ray (3, 3, 1 values). */
void orc_trace_flights_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, int64_t n,
                          orc_flight *out, int nthreads);
void orc_conditioned_eval_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls,
                             const float *q, const float *t_segment, int64_t n, float *out4);
void orc_nee_ls(const orc_medium_desc *desc, const orc_ray *rays, const float *ls, const float *normals,
                const float *points, const float *t_segment, int64_t n, float *grads, float *pdfs);

/* ProceduralNoise pieces (tests): SdfFunctions' random3, simplex3d, fbm; the scalar
 * ("var") and vector ("ls", its largest component as getKernelScale takes it) fields. */
void orc_random3(const float *p, int64_t n, float *out);
void orc_simplex3d(const float *p, int64_t n, float *out);
void orc_fbm(const double *p, int64_t n, int octaves, double *out);
void orc_noise_field(int32_t type, const double params8[8], int vec, const double *p, int64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
