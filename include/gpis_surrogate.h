/* gpis_surrogate.h — the learned flight sampler of the MI355X GPIS engine (C ABI).
 *
 * No counterpart in the reference: it stands in for the Renewal flight sampler,
 * finishGpisSegment of SparseConvolutionNoiseMedium.cpp:265-347 for 1D_sampling
 * media, and the plain method (gpis_trace_flights_1d) is both its training-data
 * source and its fallback.
 *
 * Network: 8 features -> 128 -> 128 -> 128 -> 128 -> 32 logits, GELU (tanh form),
 * BF16 weights and activations, FP32 accumulation (v_mfma_f32_16x16x32_bf16), FP32
 * master weights and Adam state.  Softmax over 32 classes = a piecewise-linear CDF
 * over s = (t - near) / (far - near): classes 0..30 are bins [edges[b], edges[b+1])
 * with uniform density inside, class 31 is "no crossing before far" (an exit).
 *
 * Features of a conditioned 1D flight (gpis_net_features_1d), in this order:
 *   0 conditioning value (previous value)      4 bound length far - near
 *   1 directional derivative (prev grad . dir)  5 mean at the origin
 *   2 SE length scale of the ray                 6 mean slope along the ray (grad mean . dir)
 *   3 variance (sigma^2)                         7 kernel type (1 = squared exponential)
 *
 * All pointers named *_dev / device arrays are device memory; calls are asynchronous
 * on `stream` unless stated.
 */
#ifndef GPIS_SURROGATE_H_
#define GPIS_SURROGATE_H_

#include "gpis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GPIS_NET_FEATURES 8
#define GPIS_NET_CLASSES 32
#define GPIS_NET_PARAMS 54816

typedef struct gpis_net gpis_net;

typedef struct gpis_net_calibration {
    float feat_mean[GPIS_NET_FEATURES];     /* standardisation: (x - mean) * inv_std */
    float feat_inv_std[GPIS_NET_FEATURES];
    float hull_lo[GPIS_NET_FEATURES];       /* training hull (axis-aligned box of raw features) */
    float hull_hi[GPIS_NET_FEATURES];
    float edges[GPIS_NET_CLASSES];          /* 31 bins over s in [0, 1]: edges[0] = 0 ... edges[31] = 1,
                                               strictly increasing except edges[1] == 0, which makes
                                               class 0 an atom at s = 0 (crossings at the origin) */
    float entropy_max;                      /* fallback above this predicted entropy (nats) */
} gpis_net_calibration;

/* Network with He-uniform weights drawn from `seed`, zero biases, zero Adam state,
 * on the current HIP device. */
gpis_status gpis_net_create(uint64_t seed, gpis_net **out);
gpis_status gpis_net_destroy(gpis_net *net);
const char *gpis_net_last_error(const gpis_net *net);

/* Flat FP32 parameters, layer after layer: W[in][out] row-major, then b[out]
 * (GPIS_NET_PARAMS values, host memory; synchronous).  Setting them resets Adam. */
gpis_status gpis_net_set_params(gpis_net *net, const float *params, int64_t count);
gpis_status gpis_net_get_params(gpis_net *net, float *params, int64_t count);
gpis_status gpis_net_set_calibration(gpis_net *net, const gpis_net_calibration *cal);

/* Features of each ray of a 1D_sampling medium (n x 8 floats, device). */
gpis_status gpis_net_features_1d(gpis_ctx *ctx, const gpis_ray_soa *rays, int64_t n, float *feats, void *stream);
/* Training class of each plain-method flight: its bin, 31 for an exit, -1 for an
 * invalid request (ignored by training). */
gpis_status gpis_net_classes_1d(gpis_net *net, const gpis_ray_soa *rays, const gpis_reply_soa *reply,
                                int64_t n, int32_t *classes, void *stream);

/* Gradient of the mean cross-entropy over a batch (batch a multiple of 64), without
 * updating: grad (GPIS_NET_PARAMS floats, device), loss_sum (1 float, device, the
 * summed -log p; may be NULL). */
gpis_status gpis_net_gradient(gpis_net *net, const float *feats, const int32_t *classes, int64_t batch,
                              float *grad, float *loss_sum, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on a batch: fused forward + backward,
 * weight gradients, update, BF16 repack. */
gpis_status gpis_net_train_step(gpis_net *net, const float *feats, const int32_t *classes, int64_t batch,
                                float lr, float *loss_sum, void *stream);
/* Polyak averaging of the Adam iterates: with beta in (0, 1) every later gpis_net_train_step
 * also updates avg = beta * avg + (1 - beta) * params, starting from the current
 * parameters; beta 0 stops it.  gpis_net_use_average makes the average the network's
 * parameters (and stops averaging).  The learning rate stays the caller's; the average
 * removes the iterates' step noise from the served network.  gpis_net_set_params stops
 * averaging. */
gpis_status gpis_net_set_averaging(gpis_net *net, float beta);
gpis_status gpis_net_use_average(gpis_net *net);
/* Class probabilities (n x 32, device) of raw features (n x 8), standardised inside
 * with the calibration. */
gpis_status gpis_net_forward(gpis_net *net, const float *feats, int64_t n, float *probs, void *stream);

/* Surrogate flights for a 1D_sampling medium (Renewal or Renewal+, stationary
 * single-level SE, sigma > 0; others: GPIS_UNSUPPORTED): each conditioned ray inside the
 * training hull and below the entropy bound gets t sampled from the network's CDF with
 * u = its jitter, and a complete sampleDistance reply (evals 0): the endpoint gradient
 * and resumeDistance's rules (GaussianProcessMedium.cpp:300-330: GPIS_HIT, or GPIS_FAIL
 * with GPIS_FAIL_OUTWARD for an up-crossing, or GPIS_EXIT with t = far and the field
 * value there), drawn from the law of the flight's conditioned 1D process (the
 * derivative along the ray from its Rice law at f(t) = 0, the crossing's sense from the
 * value at the first lattice point, the transverse gradient the plain method's own draw
 * of the normal seed).  Every other ray is traced by the plain method in the same call.
 * n_fallback (host, may be NULL) receives how many rays fell back; asking for it
 * synchronises the stream. */
gpis_status gpis_surrogate_flights_1d(gpis_ctx *ctx, gpis_net *net, const gpis_ray_soa *rays, int64_t n,
                                      const gpis_reply_soa *reply, uint32_t *n_fallback, void *stream);

/* gpis_surrogate_flights_1d without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same inputs, outputs and kernels as
 * gpis_surrogate_flights_1d (replies identical bit for bit), but the ray features, the
 * fallback list and its counters live in a fixed device array of the library (9 MiB per
 * GPU, held until the process ends) instead of buffers grown on demand.  At most 262,144
 * rays per call (larger n: GPIS_INVALID_ARGUMENT); ctx, net and the current device must be
 * the same GPU.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other host threads) race.  n_fallback, when given, synchronises the
 * stream; pass NULL under graph capture. */
gpis_status gpis_surrogate_flights_1d_noalloc(gpis_ctx *ctx, gpis_net *net, const gpis_ray_soa *rays, int64_t n,
                                              const gpis_reply_soa *reply, uint32_t *n_fallback, void *stream);

#ifdef __cplusplus
}
#endif
#endif
