/*
 * gpis_hip.h — C ABI of the MI355X (gfx950) sparse-convolution GPIS flight engine.
 *
 * This boundary replaces, for one GPU, the reference's batch boundary between a
 * renderer and the `sparse_conv_noise` medium:
 *   GpRayBatchBackend::trace / submitBatch        src/core/media/GpRayBatchBackend.hpp:43-55
 *   GpRaySegmentRequest / GpRaySegmentResult       src/core/media/GpRayBatchBackend.hpp:15-35
 *   launchSe1DRaySegments (CUDA device API)        src/gpu/cuda/RaySegmentBatch.hpp:83-93
 *   SparseConvolutionNoiseMedium::intersectGP +    src/core/media/SparseConvolutionNoiseMedium.cpp:188-347
 *     GaussianProcessMedium::resumeDistance rules  src/core/media/GaussianProcessMedium.cpp:289-369
 *
 * Conventions: every call returns a gpis_status and never throws; all array
 * arguments are DEVICE pointers owned by the caller; work is enqueued on the
 * given HIP stream (NULL = default stream) and is asynchronous unless stated.
 * One context may be used from one host thread at a time.
 */
#ifndef GPIS_HIP_H_
#define GPIS_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gpis_status {
    GPIS_OK = 0,
    GPIS_INVALID_ARGUMENT = 1,
    GPIS_DEVICE_ERROR = 2,
    GPIS_UNSUPPORTED = 3,
    GPIS_OUT_OF_MEMORY = 4
} gpis_status;

/* Covariance kinds: SquaredExponentialCovariance (GPFunctions.hpp:1880-1990) and
 * ProceduralNonstationaryCovariance (GPFunctions.hpp:2459-2505). */
enum { GPIS_COV_SE = 0, GPIS_COV_PROC_NS = 1 };
/* Mean kinds: SphericalMean (GPFunctions.hpp:952-996); box SDF and plane in code;
 * TabulatedMean over a dense grid (GPFunctions.cpp:165-190), set by gpis_set_mean_grid;
 * ProceduralMean over an SdfFunctions function ("func", GPFunctions.cpp:250-298). */
enum { GPIS_MEAN_SPHERE = 0, GPIS_MEAN_BOX = 1, GPIS_MEAN_PLANE = 2, GPIS_MEAN_GRID = 3, GPIS_MEAN_PROCEDURAL = 4 };
/* SdfFunctions::Function (SdfFunctions.hpp:13-22), the "func" of a procedural mean. */
enum { GPIS_SDF_KNOB = 0, GPIS_SDF_KNOB_INNER = 1, GPIS_SDF_KNOB_OUTER = 2, GPIS_SDF_TWO_SPHERES = 3,
       GPIS_SDF_PLANE = 4, GPIS_SDF_ROUNDED_U_GROOVE = 5, GPIS_SDF_TWO_SLABS = 6, GPIS_SDF_THREE_SLABS = 7 };
/* GPCorrelationContext (GaussianProcessMedium.cpp:20-42). */
enum { GPIS_CTX_GLOBAL = 0, GPIS_CTX_RENEWAL = 1, GPIS_CTX_RENEWAL_PLUS = 2, GPIS_CTX_NONE = 3 };
/* SparseConv1DSamplingScheme (Medium.hpp:42-46). */
enum { GPIS_SCHEME_UNI = 0, GPIS_SCHEME_NEE = 1, GPIS_SCHEME_MIS = 2 };
/* Procedural length-scale fields of the nonstationary covariance. */
enum { GPIS_LS_CONST = 0, GPIS_LS_SINCOS = 1, GPIS_LS_NOISE = 2 };
enum { GPIS_VAR_CONST = 0, GPIS_VAR_NOISE = 1 };
/* ProceduralNoise "noise" types (GPFunctions.hpp:655-690) */
enum { GPIS_NOISE_BOTTOM_TOP = 0, GPIS_NOISE_LEFT_RIGHT = 1, GPIS_NOISE_FRONT_BACK = 2,
       GPIS_NOISE_BOTTOM_TOP_LEFT_RIGHT = 3, GPIS_NOISE_SANDSTONE = 4, GPIS_NOISE_RUST = 5 };

/* The sparse_conv_noise medium's parameters (SparseConvolutionNoiseMedium.cpp:102-118,
 * GaussianProcessMedium.cpp:98-127, GPFunctions.cpp:657-685). */
typedef struct gpis_field_desc {
    int32_t cov_kind;
    float sigma, length_scale, kernel_scale;   /* "sigma", "lengthScale", "localScale" */
    float aniso[3];                            /* "aniso" */
    int32_t ns_multires;                       /* "multiResolutionGrid" */
    int32_t ns_ls_field;
    float ns_ls_params[4];                     /* SINCOS: base, amplitude, frequency */
    float ns_ls_max;
    float ns_var;
    int32_t mean_kind;
    float mean_radius;
    double mean_center[3];
    double mean_half[3];
    double plane_normal[3];
    double plane_offset;
    uint32_t global_seed;                      /* "seed" */
    float impulse_density;                     /* "impulse_density" */
    int32_t single_realization;                /* "single_realization" */
    int32_t ctxt;                              /* "correlation_context" */
    int32_t one_d;                             /* "1D_sampling" */
    int32_t scheme;                            /* "1D_sampling_scheme" */
    int32_t correlation_xy;                    /* "1D_gradient_correlationXY" */
    float step_size;                           /* "step_size" */
    uint32_t min_step;                         /* "min_step" */
    int32_t max_bounce;                        /* Medium "max_bounces" */
    int32_t isotropic_3d;                      /* "isotropic_3D_sampling": 3D noise in the isotropic
                                                  ray space of each flight (stationary SE kernel) */
    /* GPIS_MEAN_PROCEDURAL (ProceduralMean::fromJson, GPFunctions.cpp:250-266) */
    int32_t sdf_function;                      /* "func", GPIS_SDF_* */
    float sdf_inverse_transform[12];           /* _invConfigTransform rows 0-2, row-major */
    float sdf_min, sdf_scale, sdf_offset;      /* "min" (-FLT_MAX), "scale" (1), "offset" (0) */
    /* proc_nonstationary fields of type "noise" (ProceduralNoise "var", ProceduralNoiseVec
     * "ls"; GPFunctions.hpp:647-830, GPFunctions.cpp:53-135): ns_var_field GPIS_VAR_NOISE
     * and/or ns_ls_field GPIS_LS_NOISE, the GPIS_NOISE_* type of each and its JSON "min",
     * "max", "start", "end", "min2", "max2", "start2", "end2"; ns_ls_max is then the ls
     * field's maxVal() */
    int32_t ns_var_field;
    int32_t ns_var_noise, ns_ls_noise;
    double ns_var_noise_params[8], ns_ls_noise_params[8];
    /* "useAnisoMtx" and "anisoMtx" of the SE covariance (GPFunctions.cpp:663-681): when
     * use_aniso_mtx is set, the local-to-world matrix is lengthScale / sqrt(2) * aniso_mtx
     * (row-major) in place of the diagonal "aniso"; served for 3D sampling (world space
     * and the isotropic ray space) with the stationary SE kernel */
    int32_t use_aniso_mtx;
    float aniso_mtx[9];
} gpis_field_desc;

/* A wavefront of flight requests in structure-of-arrays form
 * (GpRaySegmentRequest fields, GpRayBatchBackend.hpp:15-29). */
typedef struct gpis_ray_soa {
    const float *ox, *oy, *oz;        /* origin */
    const float *dx, *dy, *dz;        /* unit direction */
    const float *near_t, *far_t;
    const float *jitter;              /* the one PathSampleGenerator draw, [0,1) */
    const uint32_t *noise_seed;       /* SparseConvolutionNoiseRealization::computeSeed */
    const uint32_t *normal_seed;      /* xxhash32(pss) + xxhash32(sceneSeed) + 1 */
    const uint32_t *flags;            /* bit 0: firstScatter */
    const float *prev_value;          /* MediumState::lastVal   (read when !firstScatter) */
    const float *prev_gx, *prev_gy, *prev_gz; /* MediumState::lastAniso (idem) */
    const float *length_scale;        /* optional per-ray SE length scale (NULL: the medium's) */
} gpis_ray_soa;

/* Outcomes of one flight (DistanceStatus and sample.exited, GaussianProcessMedium.cpp:289-369). */
enum { GPIS_EXIT = 0, GPIS_HIT = 1, GPIS_FAIL = 2, GPIS_INVALID = 3 };
enum { GPIS_FAIL_NONE = 0, GPIS_FAIL_OUTWARD = 1, GPIS_FAIL_ZERO_GRAD = 2, GPIS_FAIL_NONFINITE = 3,
       GPIS_FAIL_BOUNCE_LIMIT = 4 /* the medium's max_bounces (beginDistance, GaussianProcessMedium.cpp:252-255) */ };

typedef struct gpis_reply_soa {
    float *t;                         /* sample distance; farT on exit */
    float *value;                     /* new MediumState::lastVal (0 at a hit, f(farT) on exit) */
    float *gx, *gy, *gz;              /* sample.aniso: gradient at the hit, or at farT on exit */
    uint32_t *outcome;                /* low byte GPIS_EXIT/HIT/FAIL/INVALID, next byte fail reason */
    uint32_t *evals;                  /* field evaluations spent on this flight (may be NULL) */
} gpis_reply_soa;

typedef struct gpis_counters {
    uint64_t flights, evals, steps, hits, exits, fails, invalid;
    uint64_t kernel_launches;
    uint64_t spec_evals;              /* evaluations made ahead (two points per list pass) that the
                                         flight then did not ask for; not counted in evals */
    uint64_t passes, paired;          /* 3D flight kernels: evaluation passes of the lanes, and the
                                         passes whose second point the flight used */
} gpis_counters;

typedef struct gpis_ctx gpis_ctx;

/* Create a context for one medium on the current HIP device.
 * Replaces SparseConvolutionNoiseMedium::fromJson + the realization constructor
 * (SparseConvolutionNoiseMedium.cpp:102-118, SparseConvolutionNoise.cpp:25-54). */
gpis_status gpis_create(const gpis_field_desc *desc, gpis_ctx **out);
gpis_status gpis_destroy(gpis_ctx *ctx);
/* Text of the last error on this context (never NULL). */
const char *gpis_last_error(const gpis_ctx *ctx);

/* Precompute the impulse tables of every cell in [lo, lo+dims) for `seed`
 * (used when single_realization makes the realization a pure function of the
 * cell).  Cells outside the box are generated on the fly.  Synchronous.
 * The table is a pure function of SparseConvolutionNoise.cpp:432-452's cell3D RNG. */
gpis_status gpis_set_domain(gpis_ctx *ctx, const float lo[3], const float hi[3], uint32_t seed);

/* Write the impulses of the cells [lo, lo+dims) for `seed`: out[(cell*N + k)*4 + {x,y,z,w}],
 * cells ordered x fastest.  Bit-exact with cell3D (SparseConvolutionNoise.cpp:432-452). */
gpis_status gpis_impulse_table(gpis_ctx *ctx, const int32_t lo[3], const int32_t dims[3],
                               uint32_t seed, float *out, void *stream);

/* Value and world gradient of the unconditioned field at n points (SoA in/out):
 * evaluateValue/evaluateGradient (SparseConvolutionNoise.cpp:114-140) with seed. */
gpis_status gpis_eval_points(gpis_ctx *ctx, const float *x, const float *y, const float *z,
                             int64_t n, uint32_t seed, float *value, float *gx, float *gy,
                             float *gz, void *stream);

/* Trace n independent flights, one per request: prepareGpisSegment + finishGpisSegment
 * + resumeDistance outcome rules, in one persistent launch. */
gpis_status gpis_trace_flights(gpis_ctx *ctx, const gpis_ray_soa *rays, int64_t n,
                               const gpis_reply_soa *reply, void *stream);

/* Trace n paths of up to 1 + max_bounces flights each in one persistent launch:
 * each hit continues as a cosine-weighted diffuse bounce about the normalised
 * gradient (secondary flights use the step-1 sign rule; a conditioned medium conditions
 * each on (0, the hit's gradient), prepareGpisSegment SparseConvolutionNoiseMedium.cpp:199-241),
 * bounded by the box [-box, box]^3.  Per path: number of flights, flights per outcome,
 * the first and the last flight's reply.  Bounce directions and jitters use PCG32
 * seeded by xxhash32(ray index, bounce, bounce_seed, 0) + 1 (u1, u2, jitter; local
 * direction (sqrt(u1) cos 2 pi u2, sqrt(u1) sin 2 pi u2, sqrt(1 - u1)) in TangentFrame's
 * basis, with correctly rounded square roots and divisions and a fixed sin/cos polynomial,
 * so a host restatement reproduces it).  seeds (may be NULL): flight b of path i uses
 * the realization seed of pixelSampleSegment (px, py, spp, bounce + b) and scene_seed
 * (computeSeed, SparseConvolutionNoise.cpp:57-66); NULL: every flight of a path uses
 * rays->noise_seed. */
typedef struct gpis_path_reply {
    float *t0, *gx0, *gy0, *gz0;      /* first flight: t and gradient */
    uint32_t *outcome0;
    uint32_t *flights;                /* flights traced for this path */
    uint32_t *hits;                   /* hit flights on this path */
    uint32_t *evals;                  /* field evaluations on this path (may be NULL) */
    float *t_last;                    /* last flight: t and outcome (may be NULL, both or neither) */
    uint32_t *outcome_last;
} gpis_path_reply;
typedef struct gpis_path_seeds gpis_path_seeds;
gpis_status gpis_trace_paths(gpis_ctx *ctx, const gpis_ray_soa *rays, const gpis_path_seeds *seeds, int64_t n,
                             int32_t max_bounces, uint32_t bounce_seed, float box,
                             const gpis_path_reply *reply, void *stream);

/* Next-event estimation outputs of a 1D flight (neeGrad and neePDF,
 * SparseConvolutionNoise.cpp:662-821): the gradient implied by the mirror normal that
 * reflects the ray toward the light, and the solid-angle pdf of sampling it. */
typedef struct gpis_nee_out {
    float *gx, *gy, *gz, *pdf;
} gpis_nee_out;

/* 1D_sampling media (evaluateNoise1DNormalized / conditioning1D / evaluateGradientNoise1D,
 * SparseConvolutionNoise.cpp:178-243, 379-485, 576-656): one flight per request.
 * rays->length_scale may give each ray its own SE length scale.  When `light` is
 * non-NULL, every hit also computes NEE toward the point light at light[0..2]:
 * the mirror normal h = normalize(normalize(light - p) - dir) at the scattering point
 * p = origin + t * dir, then neeGrad(h) and neePDF(h) (zeros for other outcomes). */
gpis_status gpis_trace_flights_1d(gpis_ctx *ctx, const gpis_ray_soa *rays, int64_t n,
                                  const gpis_reply_soa *reply, const float *light,
                                  const gpis_nee_out *nee, void *stream);

/* Device-resident paths of a 1D_sampling medium: the hit-time phase and Renewal+
 * chaining of PathTracer's medium branch (PathTracer.cpp:380-438) without host round
 * trips.  Path i starts with the flight of rays[i]; at every hit the kernel
 *   - forms the scattering frame of BRDFPhaseFunction::setEventIsectInfo
 *     (BRDFPhaseFunction.cpp:29-49: TangentFrame of the normalised gradient);
 *   - when `light` is non-NULL, takes volumeLightSample's light sample toward the point
 *     light at light[0..2] (TraceBase.cpp:798-839, Point::sampleDirect Point.cpp:98-106,
 *     emission 1): f = the phase's eval (mirror: MirrorBsdf::eval = neePDF of the
 *     half-vector normal, none under the UNI sampling scheme; Lambert: cos/pi), then a
 *     shadow flight over min(distance, box exit) conditioned on (0, neeGrad) (mirror) or
 *     (0, gradient) (Lambert) whose any crossing occludes (resumeTransmittance,
 *     GaussianProcessMedium.cpp:394-421); an unoccluded sample adds f / |light - p|^2;
 *   - samples the phase (MirrorBsdf::sample, MirrorBsdf.cpp:30-49, or LambertBsdf::sample,
 *     LambertBsdf.cpp:27-38, albedo 1) and traces the next flight from the hit to the
 *     exit of [-box, box]^3, conditioned on (0, gradient) (prepareGpisSegment,
 *     SparseConvolutionNoiseMedium.cpp:199-241), with pixelSampleSegment.w one higher.
 * A path ends at an exit, a failure, a refused Lambert sample or after 1 + max_bounces
 * flights (PathTracer's maxBounces = 1 + max_bounces).  Flight b of path i uses the
 * seeds of pixelSampleSegment (px, py, spp, bounce + b) and scene_seed (computeSeed
 * SparseConvolutionNoise.cpp:57-66, prepareGpisSegment :235-238): rays->noise_seed and
 * normal_seed are not read.  The path sampler's draws are PCG32 streams seeded by
 * xxhash32(i, b, bounce_seed, s) + 1: stream 0 of flight b >= 1 gives (u1, u2, jitter)
 * of its phase sample and flight, stream 1 of flight b the jitter of its shadow flight.
 * SE kernel only; rays->length_scale may give each path its own length scale. */
enum { GPIS_PHASE_MIRROR = 0, GPIS_PHASE_LAMBERT = 1 };
struct gpis_path_seeds {
    const uint32_t *px, *py, *spp, *bounce;  /* RayInfo::pixelSampleSegment of each path's first flight */
    uint32_t scene_seed;                     /* Scene::_seed */
};
typedef struct gpis_path1d_reply {
    float *t0, *gx0, *gy0, *gz0;      /* first flight: t and gradient */
    uint32_t *outcome0;
    uint32_t *flights, *hits;         /* path flights, hits among them */
    uint32_t *shadows, *visible;      /* shadow flights, unoccluded among them */
    float *nee;                       /* sum of f / |light - p|^2 over unoccluded light samples */
    float *t_last;                    /* last path flight: t and outcome */
    uint32_t *outcome_last;
    uint32_t *evals;                  /* field evaluations of the path, shadow flights included (may be NULL) */
} gpis_path1d_reply;
gpis_status gpis_trace_paths_1d(gpis_ctx *ctx, const gpis_ray_soa *rays, const gpis_path_seeds *seeds, int64_t n,
                                int32_t max_bounces, int32_t phase, uint32_t bounce_seed, float box,
                                const float *light, const gpis_path1d_reply *reply, void *stream);

/* A tabulated mean on a dense voxel grid: the fields of the reference's HostMeanGrid
 * (src/gpu/cuda/RaySegmentBatch.hpp:19-30; made from a VDB grid by makeHostMeanGrid,
 * src/gpu/VdbMeanGrid.cpp:9-32).  density(p) is the trilinear interpolation of the
 * voxels at q = world_to_index * p clamped to [clamp_min, clamp_max] (voxel indices
 * clamped to the grid); mean(p) = (density + offset) * scale in double, and its gradient is
 * TabulatedMean::dmean_da's forward difference with eps = 1e-3 (GPFunctions.cpp:178-190). */
typedef struct gpis_mean_grid {
    const float *voxels;          /* host memory: dims[0]*dims[1]*dims[2] values, x fastest */
    int32_t origin[3];            /* index of voxels[0] */
    int32_t dims[3];
    float world_to_index[12];     /* row-major 3x4 affine map */
    float clamp_min[3], clamp_max[3];
    float offset, scale;
    int32_t volume;               /* nonzero: density -> -log(max(1e-4, density)) first
                                     (TabulatedMean's "volume" option) */
} gpis_mean_grid;
/* Copies the grid to the device; required (and only allowed) for media created with
 * mean_kind = GPIS_MEAN_GRID.  Replaces TabulatedMean::loadResources (:192-194). */
gpis_status gpis_set_mean_grid(gpis_ctx *ctx, const gpis_mean_grid *grid);

/* Accumulated device counters (synchronises the context's stream). */
gpis_status gpis_get_counters(gpis_ctx *ctx, gpis_counters *out);
gpis_status gpis_reset_counters(gpis_ctx *ctx);

/* Launch-shape control (the autotuner's knobs): waves per CU of the persistent
 * flight kernel, and rays per lane batch.  0 = built-in default. */
gpis_status gpis_set_launch(gpis_ctx *ctx, int32_t waves_per_cu, int32_t block_size);
/* Batching of the persistent flight kernels' finish/bounce/refill step (defaults: 8 lanes
 * for the 3D kernel, 16 for the 1D kernel; this call sets both): a lane whose
 * flight ended waits until `batch` lanes of its wave (1..64) can be finished, bounced
 * and refilled together, or until no lane of the wave is still marching.  1 (or 0)
 * serves each lane at once; 64 refills a wave only when all its lanes are done, so
 * rays submitted in spatially coherent groups of 64 (8x8 pixel tiles) march together.
 * No reference counterpart (a scheduling knob; results do not depend on it). */
gpis_status gpis_set_refill(gpis_ctx *ctx, int32_t batch);
/* Wave-cooperative generated cells in the 3D flight kernels (1 = on, the default where it
 * applies; 0 = off): for media whose conditioned flights generate their cells (Renewal
 * and Renewal+ over the nonstationary fields, single-level or multi-resolution, at most 12
 * impulses per cell), the lanes of a wave make the cells within reach of its points
 * together in LDS, and lanes without a point help; each point's sum is the same to the bit.
 * Other media ignore it.  No reference counterpart (a scheduling knob; results do not
 * depend on it). */
gpis_status gpis_set_coop_cells(gpis_ctx *ctx, int32_t on);

/* Support lists of the impulse table (the autotuner's knobs: "sub-cells per cell edge",
 * 1..8, default 3, 0 = none; "points per list pass", 1 or 2, default 2).  For every cell of
 * the table whose neighbourhood is in the table, each of its sub^3 sub-cells lists the
 * impulses within reach of the sub-cell (one grid unit, or, when a pass serves two points,
 * one grid unit plus the march's step_size in grid units and a 1% margin, at most 1.125),
 * in cell3D's visiting order (SparseConvolutionNoise.cpp:419-452); a lane sums only
 * its query's list, and with points = 2 the 3D flight kernels evaluate the next point the
 * march already determines (next lattice point, next secant candidate, the gradient at
 * farT) in the same pass -- over the list, or, without lists (per-flight realizations, the
 * multi-resolution grid), over the generated or table cells of the block when both points
 * lie in one grid cell.  Results are the same as without lists (same support decisions,
 * same order of the sums; replies and evaluation counts unchanged).  Rebuilds the lists of
 * an existing table (synchronous); later gpis_set_domain calls build them too.  No
 * reference counterpart (a data structure of the single-realization table). */
gpis_status gpis_set_lists(gpis_ctx *ctx, int32_t sub, int32_t points);

/* Arithmetic of the 1D along-ray flights (gpis_trace_flights_1d, gpis_trace_paths_1d and
 * the surrogate's fallback flights).  GPIS_PRECISION_FAST (the default): the splats,
 * the conditioning splats and the analytic means in FP32 with the hardware exponential,
 * within the flight tolerance of the reference (hit distances to 1e-4 relative, outcomes
 * identical but at grazing crossings).  GPIS_PRECISION_EXACT: the reference's rounding
 * step by step (exponentials rounded from double, mean in double), bit-identical with
 * the CPU restatement.  Support decisions, summation order, conditioning, the secant and
 * the transverse gradient are the same in both.  No reference counterpart (the reference
 * has one arithmetic). */
enum { GPIS_PRECISION_FAST = 0, GPIS_PRECISION_EXACT = 1 };
gpis_status gpis_set_precision(gpis_ctx *ctx, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif
