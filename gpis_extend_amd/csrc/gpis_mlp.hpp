// gpis_mlp.hpp — the learned flight sampler on gfx950 BF16 MFMA: an MLP
// 8 -> 128 -> 128 -> 128 -> 128 -> 32 (tanh-form GELU) whose softmax is a piecewise-linear CDF over
// the normalised next-crossing distance (31 bins + "exit").  No counterpart in the
// reference: it stands in for the Renewal flight sampler of
// SparseConvolutionNoiseMedium.cpp:265-347 and is measured against it.
//
// Register-resident layout.  Every layer is computed transposed, D = W^T X^T, with
// v_mfma_f32_16x16x32_bf16: the weights are the A operand (16 output features x 32
// input slots per fragment) and the activations the B operand (32 input slots x 16
// batch rows).  Lane l of a wave owns batch row l%16; the accumulator gives it output
// features 16*mt + 4*(l/16) + r (r = 0..3) of every 16-feature tile mt.  Input slots
// are numbered so that those same registers are the next layer's B fragment:
//     slot 32*kc + 8*g + s  <->  feature 32*kc + 16*(s/4) + 4*g + s%4     (g = l/16)
// and the packed weight fragments follow that numbering.  Activations therefore
// never leave registers between layers; the backward pass uses the same trick with
// W (not transposed) as the A operand.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpis {
namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 8, kWidth = 128, kClasses = 32, kBins = 31, kLayers = 5;
constexpr int kParams = 8 * 128 + 128 + 3 * (128 * 128 + 128) + 128 * 32 + 32;  // 54,816

__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kIn : kWidth; }
__host__ __device__ constexpr int layer_out(int l) { return l == kLayers - 1 ? kClasses : kWidth; }
// input slots of the forward fragments (layer 0 pads 8 features to one 32-slot chunk)
__host__ __device__ constexpr int layer_kc(int l) { return l == 0 ? 1 : kWidth / 32; }
__host__ __device__ constexpr int layer_mt(int l) { return layer_out(l) / 16; }
// natural parameter offsets: W[in][out] row-major then b[out], layer after layer
__host__ __device__ constexpr int w_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += layer_in(i) * layer_out(i) + layer_out(i);
    return o;
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + layer_in(l) * layer_out(l); }
// packed forward fragments (bf16): MT x KC fragments of 64 lanes x 8 values per layer
__host__ __device__ constexpr int fwd_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += layer_mt(i) * layer_kc(i) * 512;
    return o;
}
constexpr int kFwdElems = 8 * 1 * 512 + 3 * 8 * 4 * 512 + 2 * 4 * 512;  // 57,344
// packed backward fragments of layers 1..4 (A' = W: 16 input features x 32 output slots)
__host__ __device__ constexpr int bwd_mt(int l) { return layer_in(l) / 16; }
__host__ __device__ constexpr int bwd_kc(int l) { return layer_out(l) / 32; }
__host__ __device__ constexpr int bwd_off(int l) {
    int o = 0;
    for (int i = 1; i < l; ++i) o += bwd_mt(i) * bwd_kc(i) * 512;
    return o;
}
constexpr int kBwdElems = 3 * 8 * 4 * 512 + 8 * 1 * 512;  // 53,248
// feature-major activation / gradient workspaces: rows of layer_in (layer 0 padded to 16)
__host__ __device__ constexpr int xt_rows(int l) { return l == 0 ? 16 : kWidth; }
__host__ __device__ constexpr int xt_off_rows(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += xt_rows(i);
    return o;
}
constexpr int kXtRows = 16 + 4 * kWidth;  // 528
__host__ __device__ constexpr int dzt_off_rows(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += layer_out(i);
    return o;
}
constexpr int kDztRows = 4 * kWidth + kClasses;  // 544

// slot -> feature of a 32-slot chunk (see the header)
__host__ __device__ constexpr int slot_feature(int slot) {
    return (slot & ~31) + 16 * ((slot & 7) >> 2) + 4 * ((slot >> 3) & 3) + (slot & 3);
}

struct Calibration {
    float mean[kIn], inv_std[kIn];  // feature standardisation
    float lo[kIn], hi[kIn];         // axis-aligned hull of the training features
    float edges[kBins + 1];         // bin edges over s in [0, 1]
    float entropy_max;              // fallback when the predicted entropy exceeds this
};

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// GELU in its tanh form, 0.5 z (1 + tanh(u)) with u = sqrt(2/pi) (z + 0.044715 z^3),
// written as z * sigmoid(2u): one exp2 and one reciprocal.
constexpr float kGeluC1 = -2.0f * 0.79788456080286536f * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
constexpr float kGeluC3 = kGeluC1 * 0.044715f;
__device__ __forceinline__ float gelu_sig(float z) {
    const float a = z * fmaf(kGeluC3, z * z, kGeluC1);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float s = gelu_sig(z);
    // d/dz z s(2u) = s + z s (1 - s) 2u',  2u' = 2 sqrt(2/pi) (1 + 3 * 0.044715 z^2)
    return s + z * s * (1.0f - s) * (1.5957691216057308f * fmaf(0.134145f, z * z, 1.0f));
}

// ------------------------------------------------------------------ packing --
// Forward fragment element (layer l, index e) and backward fragment element.
__global__ void pack_kernel(const float *__restrict__ theta, bf16 *__restrict__ afwd, bf16 *__restrict__ abwd) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kFwdElems) {
        int l = 0;
        while (l + 1 < kLayers && e >= fwd_off(l + 1)) ++l;
        const int q = e - fwd_off(l);
        const int s = q & 7, lane = (q >> 3) & 63, frag = q >> 9;
        const int kc = frag % layer_kc(l), mt = frag / layer_kc(l);
        const int out = 16 * mt + (lane & 15);
        const int slot = 32 * kc + 8 * (lane >> 4) + s;
        float w = 0.0f;
        if (l == 0) {
            if (slot < kIn) w = theta[w_off(0) + slot * layer_out(0) + out];
        } else {
            w = theta[w_off(l) + slot_feature(slot) * layer_out(l) + out];
        }
        afwd[e] = (bf16)w;
    }
    const int eb = e - kFwdElems;
    if (eb >= 0 && eb < kBwdElems) {
        int l = 1;
        while (l + 1 < kLayers && eb >= bwd_off(l + 1)) ++l;
        const int q = eb - bwd_off(l);
        const int s = q & 7, lane = (q >> 3) & 63, frag = q >> 9;
        const int kc = frag % bwd_kc(l), mt = frag / bwd_kc(l);
        const int in = 16 * mt + (lane & 15);
        const int out = slot_feature(32 * kc + 8 * (lane >> 4) + s);
        abwd[eb] = (bf16)theta[w_off(l) + in * layer_out(l) + out];
    }
}

// Standardised, bf16 first-layer fragment of one row (lanes g = 0 carry the 8 features).
__device__ __forceinline__ bf16x8 input_fragment(const float *__restrict__ x, const Calibration &C, bool g0) {
    bf16x8 b;
#pragma unroll
    for (int s = 0; s < 8; ++s) b[s] = (bf16)(g0 ? (x[s] - C.mean[s]) * C.inv_std[s] : 0.0f);
    return b;
}

// ------------------------------------------------- fused forward + backward --
// One wave = 16 batch rows.  Writes every layer's bf16 input (feature-major Xt), the
// hidden pre-activations (feature-major fp32 Zt, read back by the backward pass) and
// the bf16 pre-activation gradients (feature-major dZt) for the weight-gradient kernel,
// and accumulates the summed cross-entropy.  Rows with class < 0 contribute nothing.
__global__ __launch_bounds__(256) void train_fwd_bwd_kernel(const float *__restrict__ feats,
                                                            const int32_t *__restrict__ cls, int B,
                                                            const Calibration *__restrict__ calib,
                                                            const bf16 *__restrict__ afwd,
                                                            const bf16 *__restrict__ abwd,
                                                            const float *__restrict__ theta, bf16 *__restrict__ xt,
                                                            float *__restrict__ zt, bf16 *__restrict__ dzt,
                                                            float *__restrict__ loss_sum) {
    const unsigned lane = lane64();
    const int g = lane >> 4;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int row = wave * 16 + (lane & 15);
    if (wave * 16 >= B) return;  // whole waves only (B is a multiple of 64)
    const Calibration &C = *calib;

    float x[kIn];
#pragma unroll
    for (int s = 0; s < kIn; ++s) x[s] = feats[(size_t)row * kIn + s];
    bf16x8 bx[4];
    bx[0] = input_fragment(x, C, g == 0);
    bx[1] = bx[2] = bx[3] = bx[0];
    // Xt of layer 0: features 0..7 from lanes g = 0, zero padding 8..15 from lanes g = 1
    if (g == 0) {
#pragma unroll
        for (int s = 0; s < 8; ++s) xt[(size_t)s * B + row] = bx[0][s];
    } else if (g == 1) {
#pragma unroll
        for (int s = 0; s < 8; ++s) xt[(size_t)(8 + s) * B + row] = (bf16)0.0f;
    }

#pragma unroll 1
    for (int l = 0; l < 4; ++l) {
        const bf16 *A = afwd + fwd_off(l);
        const int KC = layer_kc(l);
        const float *bias = theta + b_off(l);
        float *Z = zt + (size_t)l * kWidth * B;
        bf16 *X = xt + (size_t)xt_off_rows(l + 1) * B;
        bf16x8 nb[4];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                if (kc < KC) {
                    const bf16x8 a = *(const bf16x8 *)(A + ((mt * KC + kc) * 64 + lane) * 8);
                    acc = mfma(a, bx[kc], acc);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int f = 16 * mt + 4 * g + r;
                const float zz = acc[r] + bias[f];
                Z[(size_t)f * B + row] = zz;
                const bf16 y = (bf16)gelu(zz);
                nb[mt >> 1][(mt & 1) * 4 + r] = y;
                X[(size_t)f * B + row] = y;
            }
        }
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) bx[kc] = nb[kc];
    }
    // output layer: logits of classes 16 * mt + 4 * g + r
    float lg[2][4];
    {
        const bf16 *A = afwd + fwd_off(4);
        const float *bias = theta + b_off(4);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                const bf16x8 a = *(const bf16x8 *)(A + ((mt * 4 + kc) * 64 + lane) * 8);
                acc = mfma(a, bx[kc], acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) lg[mt][r] = acc[r] + bias[16 * mt + 4 * g + r];
        }
    }
    // softmax over the 32 classes held by the 4 lanes of this row
    float m = lg[0][0];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) m = fmaxf(m, lg[mt][r]);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    float e[2][4], se = 0.f;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            e[mt][r] = __expf(lg[mt][r] - m);
            se += e[mt][r];
        }
    se += __shfl_xor(se, 16);
    se += __shfl_xor(se, 32);
    const int target = cls[row];
    const float inv = 1.0f / se;
    float lrow = 0.f;
    bf16x8 bz[4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = 16 * mt + 4 * g + r;
            float p = e[mt][r] * inv;
            float d = target < 0 ? 0.f : p - (c == target ? 1.f : 0.f);
            if (c == target) lrow = -(lg[mt][r] - m - __logf(se));
            const bf16 db = (bf16)d;
            bz[0][mt * 4 + r] = db;
            dzt[(size_t)(dzt_off_rows(4) + c) * B + row] = db;
        }
    bz[1] = bz[2] = bz[3] = bz[0];
    // backward through layers 4..1: dY_{l-1} = W_l dZ_l, dZ_{l-1} = dY_{l-1} * gelu'(Z_{l-1})
#pragma unroll 1
    for (int l = 4; l >= 1; --l) {
        const bf16 *A = abwd + bwd_off(l);
        const int KC = bwd_kc(l);
        const float *Z = zt + (size_t)(l - 1) * kWidth * B;
        bf16 *D = dzt + (size_t)dzt_off_rows(l - 1) * B;
        bf16x8 nb[4];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                if (kc < KC) {
                    const bf16x8 a = *(const bf16x8 *)(A + ((mt * KC + kc) * 64 + lane) * 8);
                    acc = mfma(a, bz[kc], acc);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int f = 16 * mt + 4 * g + r;
                const bf16 d = (bf16)(acc[r] * gelu_grad(Z[(size_t)f * B + row]));
                nb[mt >> 1][(mt & 1) * 4 + r] = d;
                D[(size_t)f * B + row] = d;
            }
        }
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) bz[kc] = nb[kc];
    }
    // loss: the lane holding the target class carries the row's -log p
    float wsum = lrow;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) atomicAdd(loss_sum, wsum);
}

// ------------------------------------------------------- weight gradients --
// Partial dW_l[f][o] = sum over a chunk of rows of Xt_l[f][row] dZt_l[o][row] and
// db_l[o] = sum dZt_l[o][row]; one block per (chunk, layer).  Written to
// part[chunk][param] in the natural parameter layout.
__global__ __launch_bounds__(256) void wgrad_kernel(const bf16 *__restrict__ xt, const bf16 *__restrict__ dzt, int B,
                                                    int chunk_rows, float *__restrict__ part) {
    const int chunk = blockIdx.x, l = blockIdx.y;
    const unsigned lane = lane64();
    const int g = lane >> 4, wave = threadIdx.x >> 6;
    const int F = layer_in(l), O = layer_out(l);
    const int MF = xt_rows(l) / 16, MO = O / 16;
    const bf16 *X = xt + (size_t)xt_off_rows(l) * B;
    const bf16 *D = dzt + (size_t)dzt_off_rows(l) * B;
    float *P = part + (size_t)chunk * kParams;
    const int r0 = chunk * chunk_rows;
    f32x4 acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = {0.f, 0.f, 0.f, 0.f};
    const int ntiles = MF * MO;
    for (int k0 = r0; k0 < r0 + chunk_rows; k0 += 32) {
        const int kr = k0 + 8 * g;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int t = wave + 4 * j;
            if (t < ntiles) {
                const int mf = t / MO, mo = t % MO;
                const bf16x8 a = *(const bf16x8 *)(X + (size_t)(16 * mf + (lane & 15)) * B + kr);
                const bf16x8 b = *(const bf16x8 *)(D + (size_t)(16 * mo + (lane & 15)) * B + kr);
                acc[j] = mfma(a, b, acc[j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int t = wave + 4 * j;
        if (t < ntiles) {
            const int mf = t / MO, mo = t % MO;
            const int o = 16 * mo + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int f = 16 * mf + 4 * g + r;
                if (f < F) P[w_off(l) + f * O + o] = acc[j][r];
            }
        }
    }
    // bias gradients
    for (int o = threadIdx.x; o < O; o += blockDim.x) {
        float s = 0.f;
        const bf16 *d = D + (size_t)o * B + r0;
        for (int i = 0; i < chunk_rows; ++i) s += (float)d[i];
        P[b_off(l) + o] = s;
    }
}

// Sum of the partial gradients scaled by 1/B (optionally written out), then Adam.
// avg (may be null): an exponential moving average of the parameters after each update,
// avg = beta avg + (1 - beta) theta (Polyak averaging of the Adam iterates).
__global__ void adam_kernel(const float *__restrict__ part, int chunks, float inv_b, float *__restrict__ theta,
                            float *__restrict__ m, float *__restrict__ v, float *__restrict__ grad_out, float lr,
                            float bc1, float bc2, int update, float *__restrict__ avg, float avg_beta) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float gsum = 0.f;
    for (int c = 0; c < chunks; ++c) gsum += part[(size_t)c * kParams + p];
    const float gr = gsum * inv_b;
    if (grad_out) grad_out[p] = gr;
    if (!update) return;
    const float mm = 0.9f * m[p] + 0.1f * gr;
    const float vv = 0.999f * v[p] + 0.001f * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float th = theta[p] - lr * (mm * bc1) / (sqrtf(vv * bc2) + 1e-8f);
    theta[p] = th;
    if (avg) avg[p] = avg_beta * avg[p] + (1.0f - avg_beta) * th;
}

// ---------------------------------------------------------------- inference --
// Outcome word of a served row until its endpoint is completed: this bit and the class.
constexpr uint32_t kServedPending = 0x80000000u;

struct InferOut {
    float *t;
    float *value;
    float *gx, *gy, *gz;
    uint32_t *outcome;
    uint32_t *evals;
    float *probs;          // optional [n][32] class probabilities (tests)
    uint32_t *fallback;    // compacted indices of rays left to the plain method
    uint32_t *n_fallback;  // their count
};

#ifndef GPIS_INFER_WAVES
#define GPIS_INFER_WAVES 12
#endif
// waves per block (3 per SIMD: the weights' 112 KB of LDS allow one block per CU, and the
// epilogue's dependent MFMA, exponential and reciprocal chains need the third wave to hide
// their latency; 2 per SIMD measured 8% slower, 4 no faster)
constexpr int kInferWaves = GPIS_INFER_WAVES;
constexpr int kNT = 2;          // 16-row tiles per wave and pass (weight-fragment reuse)

// The persistent inference kernel: features -> standardise -> MLP -> CDF -> a sampled
// normalised distance (u = the flight's jitter, the one sampler draw a flight consumes).
// Rays outside the hull, above the entropy bound, unconditioned or non-finite go to
// the fallback list instead.
__global__ __launch_bounds__(kInferWaves * 64) void infer_kernel(
    const float *__restrict__ feats, const float *__restrict__ near_t, const float *__restrict__ far_t,
    const float *__restrict__ jitter, const uint32_t *__restrict__ flags, int n,
    const Calibration *__restrict__ calib, const bf16 *__restrict__ afwd, const float *__restrict__ theta,
    InferOut out) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    bf16 *W = (bf16 *)smem;                                           // kFwdElems
    float *bias = (float *)(smem + kFwdElems * sizeof(bf16));          // 4 * 128 + 32
    for (int i = threadIdx.x; i < kFwdElems / 8; i += blockDim.x) ((bf16x8 *)W)[i] = ((const bf16x8 *)afwd)[i];
    for (int i = threadIdx.x; i < 4 * kWidth + kClasses; i += blockDim.x) {
        const int l = i < 4 * kWidth ? i / kWidth : 4;
        const int f = i < 4 * kWidth ? i % kWidth : i - 4 * kWidth;
        bias[i] = theta[b_off(l) + f];
    }
    __syncthreads();
    const Calibration &C = *calib;
    // without the ray arrays (jitter, bounds, flags) only the probabilities are produced
    const bool sample = jitter && near_t && far_t && out.t && out.fallback && out.n_fallback;
    const unsigned lane = lane64();
    const int g = lane >> 4;
    const int wave = threadIdx.x >> 6;
    const int nwaves = gridDim.x * kInferWaves;
    const int ngroups = (n + 16 * kNT - 1) / (16 * kNT);

    for (int grp = blockIdx.x * kInferWaves + wave; grp < ngroups; grp += nwaves) {
        bf16x8 bx[kNT][4];
        bool fb[kNT];
        int rowv[kNT];
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) {
            const int row = grp * 16 * kNT + nt * 16 + (lane & 15);
            rowv[nt] = row;
            float x[kIn];
            bool ok = row < n;
#pragma unroll
            for (int s = 0; s < kIn; ++s) {
                x[s] = ok ? feats[(size_t)row * kIn + s] : 0.f;
                ok = ok && isfinite(x[s]) && x[s] >= C.lo[s] && x[s] <= C.hi[s];
            }
            if (flags && row < n && (flags[row] & 1u)) ok = false;  // unconditioned flights use the plain method
            fb[nt] = !ok;
            bx[nt][0] = input_fragment(x, C, g == 0);
        }
        // hidden layers, two output tiles (one next-layer chunk) at a time
#pragma unroll 1
        for (int l = 0; l < 4; ++l) {
            const bf16 *A = W + fwd_off(l);
            const int KC = layer_kc(l);
            bf16x8 nb[kNT][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                // the accumulators start at the bias of their four output features
                const f32x4 b0 = *(const f32x4 *)(bias + l * kWidth + 16 * (2 * c) + 4 * g);
                const f32x4 b1 = *(const f32x4 *)(bias + l * kWidth + 16 * (2 * c + 1) + 4 * g);
                f32x4 acc[kNT][2];
#pragma unroll
                for (int nt = 0; nt < kNT; ++nt) {
                    acc[nt][0] = b0;
                    acc[nt][1] = b1;
                }
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    if (kc < KC) {
                        const bf16x8 a0 = *(const bf16x8 *)(A + (((2 * c) * KC + kc) * 64 + lane) * 8);
                        const bf16x8 a1 = *(const bf16x8 *)(A + (((2 * c + 1) * KC + kc) * 64 + lane) * 8);
#pragma unroll
                        for (int nt = 0; nt < kNT; ++nt) {
                            acc[nt][0] = mfma(a0, bx[nt][kc], acc[nt][0]);
                            acc[nt][1] = mfma(a1, bx[nt][kc], acc[nt][1]);
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int nt = 0; nt < kNT; ++nt) nb[nt][c][h * 4 + r] = (bf16)gelu(acc[nt][h][r]);
            }
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) bx[nt][kc] = nb[nt][kc];
        }
        // output layer and the CDF epilogue, one 16-row tile at a time
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) {
            float lg[2][4];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const bf16x8 a = *(const bf16x8 *)(W + fwd_off(4) + ((mt * 4 + kc) * 64 + lane) * 8);
                    acc = mfma(a, bx[nt][kc], acc);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) lg[mt][r] = acc[r] + bias[4 * kWidth + 16 * mt + 4 * g + r];
            }
            float m = lg[0][0];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) m = fmaxf(m, lg[mt][r]);
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            float e[2][4], blk[2] = {0.f, 0.f}, ent = 0.f;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    e[mt][r] = __expf(lg[mt][r] - m);
                    blk[mt] += e[mt][r];
                    ent += e[mt][r] * (lg[mt][r] - m);
                }
            // block sums of the four lanes of this row, in class order
            const unsigned base = lane & 15;
            float lo4[4], hi4[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                lo4[q] = __shfl(blk[0], base + 16 * q);
                hi4[q] = __shfl(blk[1], base + 16 * q);
            }
            ent += __shfl_xor(ent, 16);
            ent += __shfl_xor(ent, 32);
            float Z = 0.f, before_lo = 0.f, before_hi = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                Z += lo4[q];
                if (q < g) before_lo += lo4[q];
            }
            float lo_total = Z;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                Z += hi4[q];
                if (q < g) before_hi += hi4[q];
            }
            const float H = __logf(Z) - ent / Z;  // entropy in nats
            const int row = rowv[nt];
            bool fallback = fb[nt] || !(H <= C.entropy_max);
            if (out.probs && row < n) {
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) out.probs[(size_t)row * kClasses + 16 * mt + 4 * g + r] = e[mt][r] / Z;
            }
            if (sample && row < n && !fallback) {
                const float u = jitter[row] * Z;  // compare in unnormalised mass
                // this lane's two class blocks start at these cumulative masses
                const float start[2] = {before_lo, lo_total + before_hi};
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const bool last = (mt == 1 && g == 3);
                    if (u >= start[mt] && (u < start[mt] + blk[mt] || last)) {
                        float cum = start[mt];
                        int cls = 16 * mt + 4 * g + 3;
                        float pc = e[mt][3], c0 = cum + e[mt][0] + e[mt][1] + e[mt][2];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            if (u < cum + e[mt][r]) {
                                cls = 16 * mt + 4 * g + r;
                                pc = e[mt][r];
                                c0 = cum;
                                break;
                            }
                            cum += e[mt][r];
                        }
                        const float nt0 = near_t[row], ft = far_t[row];
                        float t;
                        if (cls >= kBins) {
                            t = ft;  // exit
                        } else {
                            float fr = pc > 0.f ? (u - c0) / pc : 0.f;
                            fr = fminf(fmaxf(fr, 0.f), 1.f);
                            const float s = C.edges[cls] + fr * (C.edges[cls + 1] - C.edges[cls]);
                            t = nt0 + s * (ft - nt0);
                        }
                        // the distance and the class; the endpoint (gradient, outcome, the
                        // value at an exit) is completed by served_finish_kernel
                        out.t[row] = t;
                        out.outcome[row] = kServedPending | (uint32_t)cls;
                        if (out.evals) out.evals[row] = 0u;
                    }
                }
            }
            // compaction of the fallback rows: one atomic per wave, ranks by ballot; their
            // outcome word holds no pending bit until the plain method replies
            const bool mine = sample && fallback && row < n && g == 0;
            if (mine) out.outcome[row] = 3u;  // GPIS_INVALID
            const unsigned long long mask = __ballot(mine);
            if (mask) {
                const unsigned cnt = (unsigned)__popcll(mask);
                const unsigned leader = (unsigned)(__ffsll((long long)mask) - 1);
                unsigned basei = 0;
                if (lane == leader) basei = atomicAdd(out.n_fallback, cnt);
                basei = __shfl(basei, leader);
                if (mine) {
                    const unsigned rk = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                    out.fallback[basei + rk] = (uint32_t)row;
                }
            }
        }
    }
}

}  // namespace mlp
}  // namespace gpis
