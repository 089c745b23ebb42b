// gpis_field.hpp — FP32 field evaluation on gfx950: sparse-convolution noise value
// and world gradient summed over the 27 neighbour cells, multi-resolution blend,
// and Renewal/Renewal+ conditioning splats.
//
// Reference behaviour restated (SparseConvolutionNoise.cpp):
//   noise3D / cell3D                      :419-452
//   evaluateNoise3DNormalized             :308-317
//   evaluateNoise3D (world space)         :142-153
//   kernelScaleLevelRatio                 :101-111
//   evaluateValue / evaluateGradient      :114-140
// SE kernels and normalisation: GPFunctions.cpp:555-580, 706-843.
#pragma once

#include "gpis_1d.hpp"
#include "gpis_device.hpp"

#ifndef GPIS_PREFETCH_CELLS
#define GPIS_PREFETCH_CELLS 0  // double-buffer the next cell's impulses (costs 6 float4 registers)
#endif

namespace gpis {

// Per-flight realization state: the seed and the conditioning coefficients.
struct RealState {
    uint32_t seed;
    int cond;            // conditioning splats active (non-zero coefficients)
    int coop;            // generated cells made by the wave together: 0 off, 1 on, 2 on and
                         // this lane has no point of its own (it only makes cells for others)
    float ox, oy, oz;    // conditioning origin (coeff_3D.ray_origin)
    float a;             // coeff_3D.value_scale
    float bx, by, bz;    // coeff_3D.gradient_scale
};

// The support test |to|^2 < 1 of cell3D (SparseConvolutionNoise.cpp:444-448), rounded
// exactly as Vec::lengthSq: ((x*x + y*y) + z*z) with every operation rounded and no
// contraction, so an impulse at the truncation radius falls on the same side as in
// the reference.
__device__ __forceinline__ float support_r2(float tx, float ty, float tz) {
#pragma clang fp contract(off)
    float xx = tx * tx;
    float yy = ty * ty;
    float zz = tz * tz;
    return (xx + yy) + zz;
}

// One impulse-query pair in grid units: local point (lx,ly,lz), impulse (ix,iy,iz,w);
// the splat exp(-sum_i c_i to_i^2) is splattingKernel3DVal in grid units.  With an
// isotropic kernel the exponent is c * r2 (c pre-scaled by log2(e) in c0 for exp2).
template <bool ISO>
__device__ __forceinline__ void pair3(float lx, float ly, float lz, float4 im, float c0, float c1, float c2,
                                      float &v, float &ax, float &ay, float &az) {
    float tx = lx - im.x, ty = ly - im.y, tz = lz - im.z;
    float r2 = support_r2(tx, ty, tz);
    float e;
    if (ISO) e = hw_exp2(c0 * r2);
    else e = hw_exp2(c0 * (tx * tx) + c1 * (ty * ty) + c2 * (tz * tz));
    float we = (r2 < 1.0f) ? im.w * e : 0.0f;
    v += we;
    ax += we * tx;
    ay += we * ty;
    az += we * tz;
}

// The same pair with an impulse from the packed table: (x with the weight in its
// sign bit, y, z).  |x| is an operand modifier and the weight is applied with a
// bit insert, so the packed form costs no extra instruction.
// The splat runs only when some lane of the wave has the impulse in support: a
// wave-uniform branch on the ballot of the support test.  The lanes it skips would add
// exact zeros, so the sums are the same bit for bit; a wave whose lanes query nearby points
// (camera rays of one pixel tile) skips most of the 27 cells' impulses.
template <bool ISO>
__device__ __forceinline__ void pair3p(float lx, float ly, float lz, float xs, float iy, float iz, float c0, float c1,
                                       float c2, float &v, float &ax, float &ay, float &az) {
    float tx = lx - fabsf(xs), ty = ly - iy, tz = lz - iz;
    float r2 = support_r2(tx, ty, tz);
    const bool in = r2 < 1.0f;
    if (wave_any(in)) {
        float e;
        if (ISO) e = hw_exp2(c0 * r2);
        else e = hw_exp2(c0 * (tx * tx) + c1 * (ty * ty) + c2 * (tz * tz));
        float we = in ? copysignf(e, xs) : 0.0f;
        v += we;
        ax += we * tx;
        ay += we * ty;
        az += we * tz;
    }
}

// A cell at offset d in {-1,0,1} from the query's cell holds impulses whose |to| along
// an axis is at least rx (d=-1), 0 (d=0) or 1-rx (d=+1).  When the sum of these squared
// bounds exceeds 1 (with a margin far above the rounding of to), no impulse of the cell
// can pass the support test, so the cell is skipped without changing the result.
__device__ __forceinline__ float axis_bound2(float r, int d) {
    float b = d == 0 ? r : (d == 1 ? 0.0f : 1.0f - r);
    return b * b;
}
constexpr float kCullR2 = 1.0001f;

// Impulses of one cell from the generator (cell3D's RNG sequence, bit-exact):
// set_state(xxhash32(k, j, i, seed) + 1), then per impulse z, y, x, weight.
template <int NI, bool ISO>
__device__ __forceinline__ void cell_generated(int cx, int cy, int cz, int n, uint32_t seed, float lx, float ly,
                                               float lz, float c0, float c1, float c2, float &v, float &ax, float &ay,
                                               float &az) {
    Pcg g;
    g.set((uint64_t)(uint32_t)(xxh4((uint32_t)cz, (uint32_t)cy, (uint32_t)cx, seed) + 1u));
    const int N = NI > 0 ? NI : n;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        float4 im;
        im.z = u01(g.next());
        im.y = u01(g.next());
        im.x = u01(g.next());
        im.w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
        pair3<ISO>(lx, ly, lz, im, c0, c1, c2, v, ax, ay, az);
    }
}

// The generated cells of a wave's queries made by all of its lanes together.  Each lane's
// cells within reach (the cull of noise3d_raw), in the cell order of the 3x3x3 loop, are its
// jobs.  Jobs are packed column by column — the first job of every lane that has one, then
// the second, ... — and a round takes as many whole columns as the wave has active lanes:
// the active lanes make one job each (the impulses of one cell, from cell3D's stream) into
// the wave's LDS, and then every lane splats its jobs of those columns in its order.  Each
// lane's sum is therefore the sequential loop's to the bit.  A wave with many lanes holding
// points runs the plain per-lane loop instead (its columns are full, so making the cells
// apart costs only the LDS round trip); a wave with few (the end of a launch, where a few
// long chains keep their waves alive) makes their cells in parallel, with the lanes that
// have no point of their own helping.  LDS per wave: 64 job records, then the impulses
// k-major (impulse k of slot s at 64 k + s), 64 cells of n impulses.
struct CoopJob { int32_t x, y, z; uint32_t seed; };
constexpr int kCoopMaxN = 12;     // impulses per cell at most: a block of 256 lanes then holds < 64 KB
#ifndef GPIS_COOP_MAX_LIVE_LANES
#define GPIS_COOP_MAX_LIVE_LANES 32
#endif
// lanes with points above which the wave runs the per-lane loop
constexpr int kCoopMaxLive = GPIS_COOP_MAX_LIVE_LANES;
// float4 slots of one wave's part of the launch's dynamic LDS
__host__ __device__ constexpr int coop_wave_slots(int n) { return 64 + 64 * n; }
#ifdef GPIS_HOST_CHECK
inline float4 *host_coop_lds = nullptr;
inline int host_coop_max_live = kCoopMaxLive;  // the checks run both choices
static inline float4 *coop_wave_lds(int) { return host_coop_lds; }
#define GPIS_COOP_MAX_LIVE host_coop_max_live
#else
extern __shared__ float4 gpis_coop_lds[];
__device__ __forceinline__ float4 *coop_wave_lds(int n) {
    return gpis_coop_lds + (threadIdx.x >> 6) * coop_wave_slots(n);
}
#define GPIS_COOP_MAX_LIVE kCoopMaxLive
#endif

template <int NI, bool ISO>
__device__ __forceinline__ void cells_coop(float4 *lds, int ix, int iy, int iz, float rx, float ry, float rz,
                                           int n, uint32_t seed, bool idle, float c0, float c1, float c2,
                                           float &v, float &ax, float &ay, float &az) {
    const int N = NI > 0 ? NI : n;
    uint32_t m = 0;  // bit j: cell (j / 9 - 1, j / 3 % 3 - 1, j % 3 - 1) is within reach
    if (!idle) {
#pragma unroll
        for (int j = 0; j < 27; ++j)
            if (axis_bound2(rx, j / 9) + axis_bound2(ry, (j / 3) % 3) + axis_bound2(rz, j % 3) < kCullR2)
                m |= 1u << j;
    }
    const int cnt = __builtin_popcount(m);
    if (popc64(wave_ballot(cnt > 0)) > GPIS_COOP_MAX_LIVE) {
        // the sequential loop, each lane on its own
        uint32_t w = m;
        while (w) {
            const int j = __builtin_ctz(w);
            w &= w - 1u;
            const int dx = j / 9 - 1, dy = (j / 3) % 3 - 1, dz = j % 3 - 1;
            cell_generated<NI, ISO>(ix + dx, iy + dy, iz + dz, N, seed, rx - (float)dx, ry - (float)dy,
                                    rz - (float)dz, c0, c1, c2, v, ax, ay, az);
        }
        return;
    }
    const unsigned me = lane64();
    const uint64_t below = me ? (~0ull >> (64u - me)) : 0ull;
    const uint64_t act = wave_ballot(true);
    const int A = popc64(act), rank = popc64(act & below);
    CoopJob *jobs = reinterpret_cast<CoopJob *>(lds);
    float4 *imps = lds + 64;
    uint32_t wgen = m, wspl = m;  // this lane's next job to record, and to splat (lowest bit)
    int i0 = 0;                   // the round's first column
#pragma unroll 1
    for (;;) {
        // record the jobs of columns i0 .. i1 - 1, as many whole columns as fit in A slots
        int total = 0, i1 = i0;
#pragma unroll 1
        for (; i1 < 27; ++i1) {
            const uint64_t col = wave_ballot(cnt > i1);
            const int ni = popc64(col);
            if (ni == 0 || total + ni > A) break;
            if (cnt > i1) {
                const int j = __builtin_ctz(wgen);
                wgen &= wgen - 1u;
                const int slot = total + popc64(col & below);
                GPIS_CHECK_INDEX(slot, A);
                CoopJob d;
                d.x = ix + j / 9 - 1; d.y = iy + (j / 3) % 3 - 1; d.z = iz + j % 3 - 1; d.seed = seed;
                jobs[slot] = d;
            }
            total += ni;
        }
        if (total == 0) break;
        wave_sync();
        if (rank < total) {
            const CoopJob d = jobs[rank];
            Pcg g;
            g.set((uint64_t)(uint32_t)(xxh4((uint32_t)d.z, (uint32_t)d.y, (uint32_t)d.x, d.seed) + 1u));
            for (int k = 0; k < N; ++k) {
                GPIS_CHECK_INDEX(64 * k + rank, 64 * N);
                float4 im;
                im.z = u01(g.next());
                im.y = u01(g.next());
                im.x = u01(g.next());
                im.w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
                imps[64 * k + rank] = im;
            }
        }
        wave_sync();
        int cb = 0;
#pragma unroll 1
        for (int i = i0; i < i1; ++i) {
            const uint64_t col = wave_ballot(cnt > i);
            if (cnt > i) {
                const int j = __builtin_ctz(wspl);
                wspl &= wspl - 1u;
                const int slot = cb + popc64(col & below);
                const float lx = rx - (float)(j / 9 - 1), ly = ry - (float)((j / 3) % 3 - 1),
                            lz = rz - (float)(j % 3 - 1);
                for (int k = 0; k < N; ++k) pair3<ISO>(lx, ly, lz, imps[64 * k + slot], c0, c1, c2, v, ax, ay, az);
            }
            cb += popc64(col);
        }
        wave_sync();  // the next round's records and cells overwrite these
        i0 = i1;
    }
}

// One fixed-point impulse of the support lists: (X, Y, Z) in 2^-23 grid units and the
// weight's float bits.
struct Imp4 { int32_t x, y, z, w; };

// The two arrays of the support lists, read through buffer resources: a load's byte offset
// is a 32-bit VGPR and the base lives in SGPRs, so no 64-bit address arithmetic runs per
// load, and a load past an array's end returns zeros without touching memory (the lists'
// loops read one group ahead).  The host check build reads the arrays through pointers, with
// every index checked, and gives the same zeros past the end.
struct ListSrc {
#ifdef GPIS_HOST_CHECK
    const uint32_t *ent; size_t ent_n;
    const int32_t *fix; size_t fix_bytes;
#else
    __amdgpu_buffer_rsrc_t ent, fix;
#endif
};
#ifndef GPIS_HOST_CHECK
typedef unsigned int gpis_u32x4 __attribute__((ext_vector_type(4)));
constexpr int kBufferWord3 = 0x00020000;  // raw buffer, 32-bit dwords (gfx9 family)
#endif
__device__ __forceinline__ ListSrc list_src(const FieldParams &P) {
    ListSrc s;
#ifdef GPIS_HOST_CHECK
    s.ent = P.lent; s.ent_n = P.lent_n;
    s.fix = P.lfix; s.fix_bytes = P.lfix_bytes;
#else
    s.ent = __builtin_amdgcn_make_buffer_rsrc((void *)P.lent, (short)0, (int)(P.lent_n * 4u), kBufferWord3);
    s.fix = __builtin_amdgcn_make_buffer_rsrc((void *)P.lfix, (short)0, (int)P.lfix_bytes, kBufferWord3);
#endif
    return s;
}
// The four entries of list group g (byte offsets into the fixed-point impulses).
__device__ __forceinline__ void load_group(const ListSrc &s, uint32_t g, uint32_t e[4]) {
#ifdef GPIS_HOST_CHECK
    for (int i = 0; i < 4; ++i) {
        const size_t k = 4 * (size_t)g + i;
        e[i] = k < s.ent_n ? s.ent[k] : 0u;
        if (k < s.ent_n) GPIS_CHECK_INDEX(k, s.ent_n);
    }
#else
    const gpis_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(s.ent, (int)(g * 16u), 0, 0);
    e[0] = q.x; e[1] = q.y; e[2] = q.z; e[3] = q.w;
#endif
}
// The fixed-point impulse at byte offset `off`.
__device__ __forceinline__ Imp4 load_imp(const ListSrc &s, uint32_t off) {
#ifdef GPIS_HOST_CHECK
    GPIS_CHECK_INDEX((size_t)off + 15, s.fix_bytes);
    const int32_t *q = (const int32_t *)((const char *)s.fix + off);
    Imp4 r = {q[0], q[1], q[2], q[3]};
#else
    const gpis_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(s.fix, (int)off, 0, 0);
    Imp4 r = {(int32_t)q.x, (int32_t)q.y, (int32_t)q.z, (int32_t)q.w};
#endif
    return r;
}
// "This lane's predicate" gathered over a wave as a lane mask, and this lane's bit of one.
__device__ __forceinline__ uint64_t lane_mask(bool p) { return wave_ballot(p); }
__device__ __forceinline__ bool own_bit(uint64_t m) { return ((m >> lane64()) & 1u) != 0; }

// The loads of one list pass, two groups per loop iteration, software-pipelined: the group
// being summed is in registers while the other group's impulses and the next group's entries
// are in flight.  Groups past the list's `ng` read the final group of the entry array, which
// holds only the far sentinel impulse, so an odd list's last iteration adds exact zeros.
// Usage: for (j = 0; j < ng; j += 2) { step0(j); sum q0; step1(j); sum q1; }
struct ListPipe {
    const ListSrc &src;
    uint32_t first, ng, pad;
    uint32_t e0[4], e1[4];
    Imp4 q0[4], q1[4];
    __device__ __forceinline__ uint32_t group(uint32_t k) const { return k < ng ? first + k : pad; }
    __device__ __forceinline__ ListPipe(const ListSrc &s, uint32_t first_, uint32_t ng_, uint32_t pad_)
        : src(s), first(first_), ng(ng_), pad(pad_) {
        load_group(src, group(0), e0);
#pragma unroll
        for (int i = 0; i < 4; ++i) q0[i] = load_imp(src, e0[i]);
        load_group(src, group(1), e1);
    }
    // before summing q0 (group j): group j + 1's impulses and group j + 2's entries in flight
    __device__ __forceinline__ void step0(uint32_t j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) q1[i] = load_imp(src, e1[i]);
        load_group(src, group(j + 2), e0);
    }
    // before summing q1 (group j + 1): group j + 2's impulses and group j + 3's entries
    __device__ __forceinline__ void step1(uint32_t j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) q0[i] = load_imp(src, e0[i]);
        load_group(src, group(j + 3), e1);
    }
};
__device__ __forceinline__ int32_t fix_coord(float g) {
#ifdef GPIS_HOST_CHECK
    return (int32_t)lrintf(g * lists::kFixScale);
#else
    return __float2int_rn(g * lists::kFixScale);
#endif
}

// The splat of one pair given t = query - impulse in 2^-23 grid units and the weight's bits.
template <bool ISO>
__device__ __forceinline__ void pair_splat(float tx, float ty, float tz, int32_t wbits, float s0, float s1, float s2,
                                           float &v, float &ax, float &ay, float &az, uint64_t &near) {
    const float r2 = tx * tx + ty * ty + tz * tz;
    const bool in = r2 < 0x1p46f;
    near |= lane_mask(fabsf(r2 - 0x1p46f) < 0x1p46f * lists::kNearRel);
    float e;
    if (ISO) e = hw_exp2(s0 * r2);
    else e = hw_exp2(s0 * (tx * tx) + s1 * (ty * ty) + s2 * (tz * tz));
    union { int32_t i; float f; } w;
    w.i = wbits;
    const float we = in ? e * w.f : 0.0f;
    v += we;
    ax += we * tx;
    ay += we * ty;
    az += we * tz;
}

// The isotropic splat of one pair given |t|^2 (r2) and t for the moments.
__device__ __forceinline__ void pair_splat_r2(float r2, float tx, float ty, float tz, int32_t wbits, float s0,
                                              float &v, float &ax, float &ay, float &az, uint64_t &near) {
    const bool in = r2 < 0x1p46f;
    near |= lane_mask(fabsf(r2 - 0x1p46f) < 0x1p46f * lists::kNearRel);
    const float e = hw_exp2(s0 * r2);
    union { int32_t i; float f; } w;
    w.i = wbits;
    const float we = in ? e * w.f : 0.0f;
    v += we;
    ax += we * tx;
    ay += we * ty;
    az += we * tz;
}

// One query-impulse pair of a support list, in 2^-23 grid units: t = query - impulse is
// the exact integer difference rounded once to float; |t|^2 < 2^46 is the support test.
// `near` collects pairs whose |t|^2 lies within a relative kNearRel of the radius (their
// decision is left to the exact path).  s0..s2 are the exponent factors scaled by 2^-46.
template <bool ISO>
__device__ __forceinline__ void pair_fix(int32_t Gx, int32_t Gy, int32_t Gz, const Imp4 &q, float s0, float s1,
                                         float s2, float &v, float &ax, float &ay, float &az, uint64_t &near) {
    const float tx = (float)(int32_t)((uint32_t)Gx - (uint32_t)q.x);
    const float ty = (float)(int32_t)((uint32_t)Gy - (uint32_t)q.y);
    const float tz = (float)(int32_t)((uint32_t)Gz - (uint32_t)q.z);
    pair_splat<ISO>(tx, ty, tz, q.w, s0, s1, s2, v, ax, ay, az, near);
}

// The 27-cell sum over the support list of the query's sub-cell (the query's cell `lc` is
// an interior cell of the table).  (gx, gy, gz): the query in grid units, (rx, ry, rz) its
// in-cell coordinates.  Returns true when a pair fell near the support radius: the caller
// then evaluates the query on the exact path.
template <bool ISO>
__device__ __forceinline__ bool list_sum(const FieldParams &P, unsigned lc, float gx, float gy, float gz, float rx,
                                         float ry, float rz, float c0, float c1, float c2, float &v, float &ax,
                                         float &ay, float &az) {
    const unsigned S = (unsigned)P.lsub;
    const unsigned sc = ((lc * S + (unsigned)lists::sub_index(rz, S)) * S + (unsigned)lists::sub_index(ry, S)) * S +
                        (unsigned)lists::sub_index(rx, S);
    GPIS_CHECK_INDEX(sc, P.lrange_n);
    const unsigned first = P.lrange[2 * sc], ng = P.lrange[2 * sc + 1];
    const int32_t Gx = fix_coord(gx), Gy = fix_coord(gy), Gz = fix_coord(gz);
    const float s0 = c0 * 0x1p-46f, s1 = c1 * 0x1p-46f, s2 = c2 * 0x1p-46f;
    uint64_t near = 0;
    v = ax = ay = az = 0.0f;
    const ListSrc src = list_src(P);
    ListPipe L(src, first, ng, P.lent_n / 4u - 1u);
#pragma unroll 1
    for (unsigned j = 0; j < ng; j += 2) {
        L.step0(j);
#pragma unroll
        for (int i = 0; i < 4; ++i) pair_fix<ISO>(Gx, Gy, Gz, L.q0[i], s0, s1, s2, v, ax, ay, az, near);
        L.step1(j);
#pragma unroll
        for (int i = 0; i < 4; ++i) pair_fix<ISO>(Gx, Gy, Gz, L.q1[i], s0, s1, s2, v, ax, ay, az, near);
    }
    ax *= 0x1p-23f;
    ay *= 0x1p-23f;
    az *= 0x1p-23f;
    return own_bit(near);
}

// The table index of cell (ix, iy, iz) when it has support lists.
__device__ __forceinline__ bool list_cell(const FieldParams &P, int ix, int iy, int iz, unsigned &lc) {
    // signed offsets: a cell below the table's low corner must fail the test, not wrap
    const int m = P.lmargin;
    const int ux = ix - P.tlo[0], uy = iy - P.tlo[1], uz = iz - P.tlo[2];
    if (!(ux >= m && uy >= m && uz >= m && ux + m < P.tdim[0] && uy + m < P.tdim[1] && uz + m < P.tdim[2]))
        return false;
    lc = ((unsigned)uz * (unsigned)P.tdim[1] + (unsigned)uy) * (unsigned)P.tdim[0] + (unsigned)ux;
    return true;
}

// Two queries a and b summed over the support list of a's sub-cell (lists of reach 1 + e,
// b within e of a): every impulse is read once for both.  Returns the two near flags.
template <bool ISO>
__device__ __forceinline__ void list_sum2(const FieldParams &P, unsigned lc, const float ga[3], const float ra[3],
                                          const float gb[3], float c0, float c1, float c2, Eval4 &A, Eval4 &B,
                                          bool &near_a, bool &near_b) {
    const unsigned S = (unsigned)P.lsub;
    const unsigned sc = ((lc * S + (unsigned)lists::sub_index(ra[2], S)) * S + (unsigned)lists::sub_index(ra[1], S)) *
                            S + (unsigned)lists::sub_index(ra[0], S);
    GPIS_CHECK_INDEX(sc, P.lrange_n);
    const unsigned first = P.lrange[2 * sc], ng = P.lrange[2 * sc + 1];
    const int32_t Ax = fix_coord(ga[0]), Ay = fix_coord(ga[1]), Az = fix_coord(ga[2]);
    const int32_t Bx = fix_coord(gb[0]), By = fix_coord(gb[1]), Bz = fix_coord(gb[2]);
    const float dbx = (float)(Bx - Ax), dby = (float)(By - Ay), dbz = (float)(Bz - Az);
    const float s0 = c0 * 0x1p-46f, s1 = c1 * 0x1p-46f, s2 = c2 * 0x1p-46f;
    uint64_t mask_a = 0, mask_b = 0;
    A.v = A.gx = A.gy = A.gz = 0.0f;
    B.v = B.gx = B.gy = B.gz = 0.0f;
    const ListSrc src = list_src(P);
    ListPipe L(src, first, ng, P.lent_n / 4u - 1u);
    // isotropic kernels: b's |t|^2 from a's, |t + d|^2 = |t|^2 + |d|^2 + 2 t.d, and b's moments
    // summed against a's t (sum w e (t + d) = sum w e t + d sum w e, the d term added once
    // after the loop).  The rounding of this |t|^2 is far inside the near-radius band, so b's
    // support decisions stay the exact path's.
    const float d2x = 2.0f * dbx, d2y = 2.0f * dby, d2z = 2.0f * dbz;
    const float dd = dbx * dbx + dby * dby + dbz * dbz;
    auto sum4 = [&](const Imp4 *q) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float tx = (float)(int32_t)((uint32_t)Ax - (uint32_t)q[i].x);
            const float ty = (float)(int32_t)((uint32_t)Ay - (uint32_t)q[i].y);
            const float tz = (float)(int32_t)((uint32_t)Az - (uint32_t)q[i].z);
            if (ISO) {
                const float r2a = tx * tx + ty * ty + tz * tz;
                pair_splat_r2(r2a, tx, ty, tz, q[i].w, s0, A.v, A.gx, A.gy, A.gz, mask_a);
                const float r2b = fmaf(tx, d2x, fmaf(ty, d2y, fmaf(tz, d2z, r2a + dd)));
                pair_splat_r2(r2b, tx, ty, tz, q[i].w, s0, B.v, B.gx, B.gy, B.gz, mask_b);
            } else {
                pair_splat<ISO>(tx, ty, tz, q[i].w, s0, s1, s2, A.v, A.gx, A.gy, A.gz, mask_a);
                // b - impulse = (a - impulse) + (b - a): integers below 2^24 in magnitude for
                // every impulse within reach, so the float sum is exact and equals the rounded
                // integer difference (far sentinel impulses round, and stay out of support)
                pair_splat<ISO>(tx + dbx, ty + dby, tz + dbz, q[i].w, s0, s1, s2, B.v, B.gx, B.gy, B.gz, mask_b);
            }
        }
    };
#pragma unroll 1
    for (unsigned j = 0; j < ng; j += 2) {
        L.step0(j);
        sum4(L.q0);
        L.step1(j);
        sum4(L.q1);
    }
    near_a = own_bit(mask_a);
    near_b = own_bit(mask_b);
    if (ISO) {
        B.gx = B.gx + dbx * B.v;
        B.gy = B.gy + dby * B.v;
        B.gz = B.gz + dbz * B.v;
    }
    A.gx *= 0x1p-23f; A.gy *= 0x1p-23f; A.gz *= 0x1p-23f;
    B.gx *= 0x1p-23f; B.gy *= 0x1p-23f; B.gz *= 0x1p-23f;
}

// Raw 27-cell sum at point p for one grid level: returns (sum w*h, sum w*h*to).
// c0..c2 are the per-axis exponent factors already scaled by -log2(e).
// When all 27 cells lie in the precomputed table the loads of cell j+1 are issued
// before cell j is summed, so one memory latency is exposed per evaluation rather
// than one per cell; otherwise each cell comes from the table or the generator.
template <int NI, bool ISO>
__device__ __forceinline__ Eval4 noise3d_raw(const FieldParams &P, float px, float py, float pz, uint32_t seed, float R,
                                             float c0, float c1, float c2, bool use_lists = true, int coop = 0) {
    float gx = px / R, gy = py / R, gz = pz / R;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float rx = gx - fx, ry = gy - fy, rz = gz - fz;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float v = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
    const int N = NI > 0 ? NI : P.N;
    bool tab = P.has_table && seed == P.tseed && coop != 2;  // a helper lane (coop 2) has no point
    // the 3x3x3 block's low cell relative to the table; signed, so that a query below the
    // table's low corner fails the test instead of wrapping around to a huge offset
    const int sbx = ix - 1 - P.tlo[0], sby = iy - 1 - P.tlo[1], sbz = iz - 1 - P.tlo[2];
    const bool inner = tab && sbx >= 0 && sby >= 0 && sbz >= 0 && sbx + 2 < P.tdim[0] && sby + 2 < P.tdim[1] &&
                       sbz + 2 < P.tdim[2];
    const unsigned bx = (unsigned)sbx, by = (unsigned)sby, bz = (unsigned)sbz;
    unsigned lc = 0;
    if (use_lists && tab && P.has_lists && list_cell(P, ix, iy, iz, lc)) {
        // the query's support list; a query with a pair at the support radius runs the
        // exact path below instead
        const bool near = list_sum<ISO>(P, lc, gx, gy, gz, rx, ry, rz, c0, c1, c2, v, ax, ay, az);
        if (!wave_any(near) || !near) {
            Eval4 r = {v, ax, ay, az};
            return r;
        }
        v = ax = ay = az = 0.0f;
    }
    if (NI > 0 && (NI * 3) % 4 == 0 && inner) {
        // fast path: the whole 3x3x3 block is in the packed table; the loads of cell
        // j+1 are in flight while cell j is summed
        constexpr int Q = NI > 0 ? NI * 3 / 4 : 1;  // float4 per cell
        const unsigned cs = NI * 3;                 // floats per cell
        const unsigned sy = (unsigned)P.tdim[0] * cs, sz = (unsigned)P.tdim[0] * (unsigned)P.tdim[1] * cs;
        const float *base = P.table + ((size_t)bz * sz + (size_t)by * sy + (size_t)bx * cs);
#if GPIS_PREFETCH_CELLS
        float4 cur[Q], nxt[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) cur[k] = reinterpret_cast<const float4 *>(base)[k];
#pragma unroll 1
        for (int j = 0; j < 27; ++j) {
            int dxi = j / 9, dyi = (j / 3) % 3, dzi = j % 3;  // the reference's dx, dy, dz order
            if (j < 26) {
                int jn = j + 1;
                const float4 *q = reinterpret_cast<const float4 *>(
                    base + ((size_t)(jn % 3) * sz + (size_t)((jn / 3) % 3) * sy + (size_t)(jn / 9) * cs));
#pragma unroll
                for (int k = 0; k < Q; ++k) nxt[k] = q[k];
            }
            float lx = rx - (float)(dxi - 1), ly = ry - (float)(dyi - 1), lz = rz - (float)(dzi - 1);
            const float *f = reinterpret_cast<const float *>(cur);
#pragma unroll
            for (int k = 0; k < (NI > 0 ? NI : 1); ++k)
                pair3p<ISO>(lx, ly, lz, f[3 * k], f[3 * k + 1], f[3 * k + 2], c0, c1, c2, v, ax, ay, az);
#pragma unroll
            for (int k = 0; k < Q; ++k) cur[k] = nxt[k];
        }
#else
        // the other resident waves hide each cell's load latency; the dz row is unrolled
        const float bz0 = axis_bound2(rz, 0), bz2 = axis_bound2(rz, 2);
#pragma unroll 1
        for (int j = 0; j < 9; ++j) {
            const int dxi = j / 3, dyi = j % 3;  // the reference's dx, dy order; dz innermost
            const float bxy = axis_bound2(rx, dxi) + axis_bound2(ry, dyi);
            if (bxy >= kCullR2) continue;
            const float *row = base + ((size_t)dyi * sy + (size_t)dxi * cs);
            const float lx = rx - (float)(dxi - 1), ly = ry - (float)(dyi - 1);
#pragma unroll
            for (int dzi = 0; dzi < 3; ++dzi) {
                const float bzz = dzi == 0 ? bz0 : (dzi == 1 ? 0.0f : bz2);
                if (bxy + bzz >= kCullR2) continue;
                GPIS_CHECK_INDEX((size_t)(row - P.table) + (size_t)dzi * sz + cs - 1, (size_t)P.table_n);
                const float4 *q = reinterpret_cast<const float4 *>(row + (size_t)dzi * sz);
                float4 cur[Q];
#pragma unroll
                for (int k = 0; k < Q; ++k) cur[k] = q[k];
                const float lz = rz - (float)(dzi - 1);
                const float *f = reinterpret_cast<const float *>(cur);
#pragma unroll
                for (int k = 0; k < (NI > 0 ? NI : 1); ++k)
                    pair3p<ISO>(lx, ly, lz, f[3 * k], f[3 * k + 1], f[3 * k + 2], c0, c1, c2, v, ax, ay, az);
            }
        }
#endif
        Eval4 r = {v, ax, ay, az};
        return r;
    }
    if (coop && !tab) {
        cells_coop<NI, ISO>(coop_wave_lds(N), ix, iy, iz, rx, ry, rz, N, seed, coop == 2, c0, c1, c2, v, ax, ay, az);
        Eval4 r = {v, ax, ay, az};
        return r;
    }
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                if (axis_bound2(rx, dx + 1) + axis_bound2(ry, dy + 1) + axis_bound2(rz, dz + 1) >= kCullR2) continue;
                int cx = ix + dx, cy = iy + dy, cz = iz + dz;
                float lx = rx - (float)dx, ly = ry - (float)dy, lz = rz - (float)dz;
                unsigned ux = (unsigned)(cx - P.tlo[0]), uy = (unsigned)(cy - P.tlo[1]), uz = (unsigned)(cz - P.tlo[2]);
                if (tab && ux < (unsigned)P.tdim[0] && uy < (unsigned)P.tdim[1] && uz < (unsigned)P.tdim[2]) {
                    const size_t qi = ((size_t)(uz * (unsigned)P.tdim[1] + uy) * (unsigned)P.tdim[0] + ux) * (3 * N);
                    GPIS_CHECK_INDEX(qi + 3 * N - 1, (size_t)P.table_n);
                    const float *q = P.table + qi;
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        pair3p<ISO>(lx, ly, lz, q[3 * k], q[3 * k + 1], q[3 * k + 2], c0, c1, c2, v, ax, ay, az);
                } else {
                    cell_generated<NI, ISO>(cx, cy, cz, N, seed, lx, ly, lz, c0, c1, c2, v, ax, ay, az);
                }
            }
        }
    }
    Eval4 r = {v, ax, ay, az};
    return r;
}

// Two queries in the same grid cell (ix, iy, iz) of one level, summed in one pass over their
// 27-cell block: each cell's impulses are read from the table or generated (cell3D's stream)
// once and splatted at both points.  Each point keeps its own cell cull and support test and
// runs over the same cells and impulses in the same order as noise3d_raw, so each sum is that
// point's single-query sum to the bit.  ra, rb: the points' in-cell coordinates.
template <int NI, bool ISO>
__device__ __forceinline__ void noise3d_cells2(const FieldParams &P, int ix, int iy, int iz, const float ra[3],
                                               const float rb[3], uint32_t seed, float c0, float c1, float c2,
                                               Eval4 &A, Eval4 &B) {
    float av = 0.f, aax = 0.f, aay = 0.f, aaz = 0.f, bv = 0.f, bax = 0.f, bay = 0.f, baz = 0.f;
    const int N = NI > 0 ? NI : P.N;
    const bool tab = P.has_table && seed == P.tseed;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                const bool na = axis_bound2(ra[0], dx + 1) + axis_bound2(ra[1], dy + 1) + axis_bound2(ra[2], dz + 1) <
                                kCullR2;
                const bool nb = axis_bound2(rb[0], dx + 1) + axis_bound2(rb[1], dy + 1) + axis_bound2(rb[2], dz + 1) <
                                kCullR2;
                if (!na && !nb) continue;
                const int cx = ix + dx, cy = iy + dy, cz = iz + dz;
                const float lax = ra[0] - (float)dx, lay = ra[1] - (float)dy, laz = ra[2] - (float)dz;
                const float lbx = rb[0] - (float)dx, lby = rb[1] - (float)dy, lbz = rb[2] - (float)dz;
                const unsigned ux = (unsigned)(cx - P.tlo[0]), uy = (unsigned)(cy - P.tlo[1]),
                               uz = (unsigned)(cz - P.tlo[2]);
                if (tab && ux < (unsigned)P.tdim[0] && uy < (unsigned)P.tdim[1] && uz < (unsigned)P.tdim[2]) {
                    const size_t qi = ((size_t)(uz * (unsigned)P.tdim[1] + uy) * (unsigned)P.tdim[0] + ux) * (3 * N);
                    GPIS_CHECK_INDEX(qi + 3 * N - 1, (size_t)P.table_n);
                    const float *q = P.table + qi;
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        if (na) pair3p<ISO>(lax, lay, laz, q[3 * k], q[3 * k + 1], q[3 * k + 2], c0, c1, c2, av, aax, aay, aaz);
                        if (nb) pair3p<ISO>(lbx, lby, lbz, q[3 * k], q[3 * k + 1], q[3 * k + 2], c0, c1, c2, bv, bax, bay, baz);
                    }
                } else {
                    Pcg g;
                    g.set((uint64_t)(uint32_t)(xxh4((uint32_t)cz, (uint32_t)cy, (uint32_t)cx, seed) + 1u));
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        float4 im;
                        im.z = u01(g.next());
                        im.y = u01(g.next());
                        im.x = u01(g.next());
                        im.w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
                        if (na) pair3<ISO>(lax, lay, laz, im, c0, c1, c2, av, aax, aay, aaz);
                        if (nb) pair3<ISO>(lbx, lby, lbz, im, c0, c1, c2, bv, bax, bay, baz);
                    }
                }
            }
        }
    }
    A.v = av; A.gx = aax; A.gy = aay; A.gz = aaz;
    B.v = bv; B.gx = bax; B.gy = bay; B.gz = baz;
}

// Two queries at once (a and, when `two`, b): when lists serve two queries and b lies within
// their extra reach of a, one pass over a's list sums both; when b shares a's grid cell and
// a has no list, one pass over the block's table or generated cells sums both; otherwise each
// is summed on its own.  Queries with a pair at the support radius are summed again on the
// exact path.
template <int NI, bool ISO>
__device__ __forceinline__ void noise3d_raw2(const FieldParams &P, const float pa[3], const float pb[3], bool two,
                                             uint32_t seed, float R, float c0, float c1, float c2, Eval4 &A,
                                             Eval4 &B) {
    // a query without a second point rides the same pass with b = a (its second sums are
    // discarded), so that the lanes of a wave stay on one path
    const float ga[3] = {pa[0] / R, pa[1] / R, pa[2] / R};
    const float gb[3] = {two ? pb[0] / R : ga[0], two ? pb[1] / R : ga[1], two ? pb[2] / R : ga[2]};
    const float fa[3] = {floorf(ga[0]), floorf(ga[1]), floorf(ga[2])};
    const float ra[3] = {ga[0] - fa[0], ga[1] - fa[1], ga[2] - fa[2]};
    const float dx = gb[0] - ga[0], dy = gb[1] - ga[1], dz = gb[2] - ga[2];
    unsigned lc = 0;
    const bool pair = P.lpair2 > 0.0f && P.has_table && seed == P.tseed && P.has_lists &&
                      dx * dx + dy * dy + dz * dz <= P.lpair2 && list_cell(P, (int)fa[0], (int)fa[1], (int)fa[2], lc);
    bool near_a = false, near_b = false;
    if (wave_any(pair)) {
        if (pair) list_sum2<ISO>(P, lc, ga, ra, gb, c0, c1, c2, A, B, near_a, near_b);
    }
    // b in a's grid cell, a without a list: one pass over the cells of the block
    unsigned lca = 0;
    const bool listed = P.has_lists && P.has_table && seed == P.tseed &&
                        list_cell(P, (int)fa[0], (int)fa[1], (int)fa[2], lca);
    const bool cells2 = two && !pair && !listed && floorf(gb[0]) == fa[0] && floorf(gb[1]) == fa[1] &&
                        floorf(gb[2]) == fa[2];
    if (wave_any(cells2)) {
        if (cells2) {
            const float rb[3] = {gb[0] - fa[0], gb[1] - fa[1], gb[2] - fa[2]};
            noise3d_cells2<NI, ISO>(P, (int)fa[0], (int)fa[1], (int)fa[2], ra, rb, seed, c0, c1, c2, A, B);
        }
    }
    // the rest one query at a time (one inlined copy of the single-query sum): queries off the
    // pair paths on their own lists or the table, near-radius queries on the exact path
#pragma unroll 1
    for (int w = 0; w < 2; ++w) {
        const bool need = !cells2 && (w == 0 ? (!pair || near_a) : (two && (!pair || near_b)));
        if (!wave_any(need)) continue;
        if (need) {
            const Eval4 r = noise3d_raw<NI, ISO>(P, w ? pb[0] : pa[0], w ? pb[1] : pa[1], w ? pb[2] : pa[2], seed, R, c0,
                                                 c1, c2, !pair);
            if (w) B = r;
            else A = r;
        }
    }
}

// Grid constants of one level with spatial scale gs (multi-resolution level
// scale, or 1 for a stationary kernel) and local scale ls:
//   R = splattingKernelRadius(false, gs)         (GPFunctions.cpp:706-719)
//   m_i = l_aniso_inv_i^2 / gs^2 / ls^2 * 0.5    (getInvCovMtx, :784-812)
//   variance = N / R^3 * pi^1.5 * prod(l_aniso) * gs^3 * ls^3  (:751-764)
constexpr float kLog2e = 1.4426950408889634f;

struct Level {
    float R, c0, c1, c2, g0, g1, g2, nf_inv;
    uint32_t seed_add;
};
__device__ __forceinline__ Level make_level(const FieldParams &P, float gs, float ls, uint32_t seed_add) {
    Level L;
    L.R = P.kernel_scale * gs * P.l_max;
    float s2 = gs * gs * ls * ls;
    float m0 = P.l_aniso_inv[0] * P.l_aniso_inv[0] / s2 * 0.5f;
    float m1 = P.l_aniso_inv[1] * P.l_aniso_inv[1] / s2 * 0.5f;
    float m2 = P.l_aniso_inv[2] * P.l_aniso_inv[2] / s2 * 0.5f;
    float R2 = L.R * L.R;
    L.c0 = -kLog2e * R2 * m0; L.c1 = -kLog2e * R2 * m1; L.c2 = -kLog2e * R2 * m2;
    L.g0 = -2.f * L.R * m0; L.g1 = -2.f * L.R * m1; L.g2 = -2.f * L.R * m2;
    float var = P.Nf / (L.R * L.R * L.R) * 5.568327996831708f * (P.l_aniso[0] * P.l_aniso[1] * P.l_aniso[2]) *
                (gs * gs * gs) * (ls * ls * ls);
    L.nf_inv = 1.0f / sqrtf(var);
    L.seed_add = seed_add;
    return L;
}

// Conditioning splats κ(p,o)·a + Σ b_i ∂_iκ(p,o) and their gradient for one level
// (splattingKernel3D / splattingKernel3DGrad with isCov, GPFunctions.cpp:555-580, 829-843).
__device__ __forceinline__ void cond_splat(const FieldParams &P, const RealState &rs, float px, float py, float pz,
                                           float gs, float ls, Eval4 &n) {
    float s2 = gs * gs * ls * ls;
    float m0 = P.l_aniso_inv[0] * P.l_aniso_inv[0] / s2 * 0.25f;
    float m1 = P.l_aniso_inv[1] * P.l_aniso_inv[1] / s2 * 0.25f;
    float m2 = P.l_aniso_inv[2] * P.l_aniso_inv[2] / s2 * 0.25f;
    float ax = px - rs.ox, ay = py - rs.oy, az = pz - rs.oz;
    float k = __expf(-(ax * ax * m0 + ay * ay * m1 + az * az * m2));
    float d0 = -2.f * ax * m0, d1 = -2.f * ay * m1, d2 = -2.f * az * m2;
    float bd = rs.bx * d0 + rs.by * d1 + rs.bz * d2;
    n.v += (rs.a + bd) * k;
    n.gx += (rs.a * d0 + bd * d0 - 2.f * m0 * rs.bx) * k;
    n.gy += (rs.a * d1 + bd * d1 - 2.f * m1 * rs.by) * k;
    n.gz += (rs.a * d2 + bd * d2 - 2.f * m2 * rs.bz) * k;
}

// One pair of the full-matrix splat (useAnisoMtx): the exponent is the quadratic form
// t^T M t in grid units with the host's coefficients fc (splattingKernel3DVal with the
// matrix getInvCovMtx, GPFunctions.cpp:784-821); the support test is the same |t|^2 < 1.
__device__ __forceinline__ void pair_full(const float fc[6], float tx, float ty, float tz, float w, float &v,
                                          float &ax, float &ay, float &az) {
    const float r2 = support_r2(tx, ty, tz);
    const float q = fc[0] * (tx * tx) + fc[1] * (ty * ty) + fc[2] * (tz * tz) + fc[3] * (tx * ty) +
                    fc[4] * (tx * tz) + fc[5] * (ty * tz);
    const float we = (r2 < 1.0f) ? w * hw_exp2(q) : 0.0f;
    v += we;
    ax += we * tx;
    ay += we * ty;
    az += we * tz;
}

// The 27-cell sum of the full-matrix splat at world point p (noise3D, SparseConvolutionNoise.cpp
// :419-452): impulses from the packed table where it holds the cell, else from the generator;
// returns (sum w e, sum w e t) in grid units.  Out of line: only the kernels instantiated with
// the generic impulse count (NI = -1) contain the call.
__device__ __noinline__ Eval4 noise3d_full(const FieldParams &P, float px, float py, float pz, uint32_t seed) {
    const float gx = px / P.R, gy = py / P.R, gz = pz / P.R;
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    const float rx = gx - fx, ry = gy - fy, rz = gz - fz;
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const bool tab = P.has_table && seed == P.tseed;
    const int N = P.N;
    float v = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                if (axis_bound2(rx, dx + 1) + axis_bound2(ry, dy + 1) + axis_bound2(rz, dz + 1) >= kCullR2) continue;
                const int cx = ix + dx, cy = iy + dy, cz = iz + dz;
                const float lx = rx - (float)dx, ly = ry - (float)dy, lz = rz - (float)dz;
                const unsigned ux = (unsigned)(cx - P.tlo[0]), uy = (unsigned)(cy - P.tlo[1]),
                               uz = (unsigned)(cz - P.tlo[2]);
                if (tab && ux < (unsigned)P.tdim[0] && uy < (unsigned)P.tdim[1] && uz < (unsigned)P.tdim[2]) {
                    const size_t qi = ((size_t)(uz * (unsigned)P.tdim[1] + uy) * (unsigned)P.tdim[0] + ux) * (3 * N);
                    GPIS_CHECK_INDEX(qi + 3 * N - 1, (size_t)P.table_n);
                    const float *q = P.table + qi;
                    for (int k = 0; k < N; ++k) {
                        const float xs = q[3 * k];
                        pair_full(P.fc, lx - fabsf(xs), ly - q[3 * k + 1], lz - q[3 * k + 2], copysignf(1.0f, xs), v,
                                  ax, ay, az);
                    }
                } else {
                    Pcg g;
                    g.set((uint64_t)(uint32_t)(xxh4((uint32_t)cz, (uint32_t)cy, (uint32_t)cx, seed) + 1u));
                    for (int k = 0; k < N; ++k) {
                        const float iz_ = u01(g.next());
                        const float iy_ = u01(g.next());
                        const float ix_ = u01(g.next());
                        const float w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
                        pair_full(P.fc, lx - ix_, ly - iy_, lz - iz_, w, v, ax, ay, az);
                    }
                }
            }
        }
    }
    Eval4 r = {v, ax, ay, az};
    return r;
}

// The normalised full-matrix noise: value and the world gradient -2 R M (sum w e t).
__device__ __forceinline__ Eval4 noise_full_normalized(const FieldParams &P, float px, float py, float pz,
                                                       uint32_t seed) {
    const Eval4 n = noise3d_full(P, px, py, pz, seed);
    const float *G = P.fg;
    Eval4 r;
    r.v = n.v * P.nf_inv;
    r.gx = (G[0] * n.gx + G[1] * n.gy + G[2] * n.gz) * P.nf_inv;
    r.gy = (G[3] * n.gx + G[4] * n.gy + G[5] * n.gz) * P.nf_inv;
    r.gz = (G[6] * n.gx + G[7] * n.gy + G[8] * n.gz) * P.nf_inv;
    return r;
}

// Conditioning splats of the full matrix: kernel exp(-a^T K a) with K = cinv / 4 (isCov),
// first derivatives d = -2 K a and second derivatives d d^T - 2 K (splattingKernel3D /
// splattingKernel3DGrad with splattingKernel3D2ndGrad, GPFunctions.cpp:555-580, 829-843).
__device__ __forceinline__ void cond_splat_full(const FieldParams &P, const RealState &rs, float px, float py,
                                                float pz, Eval4 &n) {
    const float *K = P.fk;  // 00 11 22 01 02 12
    const float ax = px - rs.ox, ay = py - rs.oy, az = pz - rs.oz;
    const float r0 = ax * K[0] + ay * K[3] + az * K[4];
    const float r1 = ax * K[3] + ay * K[1] + az * K[5];
    const float r2 = ax * K[4] + ay * K[5] + az * K[2];
    const float k = __expf(-(r0 * ax + r1 * ay + r2 * az));
    const float d0 = -2.f * r0, d1 = -2.f * r1, d2 = -2.f * r2;
    const float bd = rs.bx * d0 + rs.by * d1 + rs.bz * d2;
    const float kb0 = K[0] * rs.bx + K[3] * rs.by + K[4] * rs.bz;
    const float kb1 = K[3] * rs.bx + K[1] * rs.by + K[5] * rs.bz;
    const float kb2 = K[4] * rs.bx + K[5] * rs.by + K[2] * rs.bz;
    n.v += (rs.a + bd) * k;
    n.gx += (rs.a * d0 + bd * d0 - 2.f * kb0) * k;
    n.gy += (rs.a * d1 + bd * d1 - 2.f * kb1) * k;
    n.gz += (rs.a * d2 + bd * d2 - 2.f * kb2) * k;
}

// Kernels with the generic impulse count instantiated as NI = -1 carry the "noise" fields of
// the nonstationary covariance and the full anisotropy matrix.
template <int NI> struct HasNoise { static constexpr bool value = NI < 0; };

// getKernelScale (GPFunctions.cpp:1734-1740): the largest float component of the ls field at
// p (ls_field 0: constant; 1: the sin*cos field defined in code; 2: ProceduralNoiseVec, only
// in the kernels instantiated with NOISE: the out-of-line fbm changes a whole kernel's
// register allocation, so the other kernels never contain the call).
template <bool NOISE>
__device__ __forceinline__ float ls_field(const FieldParams &P, float x, float y, float z) {
#pragma clang fp contract(off)
    if (NOISE && P.ls_field == 2) return (float)noise_field(P.ls_noise, true, (double)x, (double)y, (double)z);
    if (P.ls_field == 1) {
        const double fr = P.ls_params[2];
        return (float)((double)P.ls_params[0] +
                       (double)P.ls_params[1] * (0.5 + 0.5 * sin(fr * (double)x) * cos(fr * (double)z)));
    }
    return P.ls_params[0];
}

// Levels of the multi-resolution ladder of base 2.5 (kernelScaleLevelRatio,
// SparseConvolutionNoise.cpp:101-111) at lateral scale lat:
//   L = floor(log(lat) / log(2.5)), scales 2.5^L and 2.5^(L+1), weights
//   (1 - fr) / sqrt(1 - 2 fr + 2 fr^2) and fr / sqrt(...), fr = (lat - 2.5^L) / (2.5^(L+1) - 2.5^L).
// The scales are exact roundings of 2.5^L (5^L and 2^L are exact in double for the clamped
// range, and the one division or power-of-two scaling rounds once), so every level's grid
// and support radius are those of the float pow in the reference.  The float logarithm is
// rounded from the double one: over every float lat in [2^-12, 2^12] its floor quotient
// equals the one of the float logf.  L is clamped to [-32, 31], the seed-offset table.
constexpr float kLogLevelBase = 0.91629076f;  // logf(2.5f)
struct LevelRatio {
    float slo, shi, wlo, whi;
    int li;
};
__device__ __forceinline__ float pow_level_base(int L) {
    const int a = L < 0 ? -L : L;
    double v = 1.0;
    for (int i = 0; i < a; ++i) v *= 5.0;
    return L >= 0 ? (float)ldexp(v, -L) : (float)(ldexp(1.0, a) / v);
}
__device__ __forceinline__ LevelRatio level_ratio(float lat) {
#pragma clang fp contract(off)
    float lvl = floorf((float)log((double)lat) / kLogLevelBase);
    lvl = fminf(fmaxf(lvl, -32.0f), 31.0f);
    LevelRatio k;
    k.li = (int)lvl;
    k.slo = pow_level_base(k.li);
    k.shi = pow_level_base(k.li + 1);
    const float fr = (lat - k.slo) / (k.shi - k.slo);
    const double fd = fr;
    const double den = sqrt(1.0 - 2.0 * fd + 2.0 * fd * fd);
    k.wlo = (float)((1.0 - fd) / den);
    k.whi = (float)(fd / den);
    return k;
}

// worldToLocal v: diag(1 / l_aniso) v, or W v for the full anisotropy matrix (each component
// left to right, as Eigen's 3x3 product).
__device__ __forceinline__ void world_to_local(const FieldParams &P, float vx, float vy, float vz, float &ix, float &iy,
                                               float &iz) {
#pragma clang fp contract(off)
    if (P.full) {
        const float *W = P.w2l;
        ix = dot3(W[0], W[1], W[2], vx, vy, vz);
        iy = dot3(W[3], W[4], W[5], vx, vy, vz);
        iz = dot3(W[6], W[7], W[8], vx, vy, vz);
    } else {
        ix = P.l_aniso_inv[0] * vx; iy = P.l_aniso_inv[1] * vy; iz = P.l_aniso_inv[2] * vz;
    }
}
// The isotropic ray space of a flight along (dx, dy, dz): the frame of
// normalize(worldToLocal d) (evaluateNoise3DIsotropicRayNormalized :349-352).
__device__ __forceinline__ Frame3 iso_ray_frame(const FieldParams &P, float dx, float dy, float dz) {
#pragma clang fp contract(off)
    float ix, iy, iz;
    world_to_local(P, dx, dy, dz, ix, iy, iz);
    normalize3(ix, iy, iz);
    return make_frame(ix, iy, iz);
}
// A world point in the isotropic ray space: frame.toLocal(worldToLocal p).
__device__ __forceinline__ void to_iso_ray(const FieldParams &P, const Frame3 &F, float px, float py, float pz,
                                           float &rx, float &ry, float &rz) {
#pragma clang fp contract(off)
    float ix, iy, iz;
    world_to_local(P, px, py, pz, ix, iy, iz);
    rx = dot3(F.tx, F.ty, F.tz, ix, iy, iz);
    ry = dot3(F.bx, F.by, F.bz, ix, iy, iz);
    rz = dot3(F.nx, F.ny, F.nz, ix, iy, iz);
}
// A gradient from the isotropic ray space to world space: localToWorld^-T frame.toGlobal g
// (worldToLocal^T g: the diagonal, or W^T for the full matrix).
__device__ __forceinline__ void from_iso_ray(const FieldParams &P, const Frame3 &F, float gx, float gy, float gz,
                                             float &wx, float &wy, float &wz) {
#pragma clang fp contract(off)
    float x = F.tx * gx + F.bx * gy;
    x += F.nx * gz;
    float y = F.ty * gx + F.by * gy;
    y += F.ny * gz;
    float z = F.tz * gx + F.bz * gy;
    z += F.nz * gz;
    if (P.full) {
        const float *W = P.w2l;
        wx = dot3(W[0], W[3], W[6], x, y, z);
        wy = dot3(W[1], W[4], W[7], x, y, z);
        wz = dot3(W[2], W[5], W[8], x, y, z);
    } else {
        wx = P.l_aniso_inv[0] * x; wy = P.l_aniso_inv[1] * y; wz = P.l_aniso_inv[2] * z;
    }
}

// One level's term of a multi-resolution evaluation (evaluateNoise3D :146-153): the level's raw
// sums normalised, its conditioning splats added, weighted by w and added to acc (the first
// level sets acc).
__device__ __forceinline__ void ml_level(const FieldParams &P, const RealState &rs, float px, float py, float pz,
                                         bool cond, float scale, float w, const Level &L, const Eval4 &n, bool first,
                                         Eval4 &acc) {
    Eval4 r = {n.v * L.nf_inv, n.gx * L.g0 * L.nf_inv, n.gy * L.g1 * L.nf_inv, n.gz * L.g2 * L.nf_inv};
    if (cond && rs.cond) cond_splat(P, rs, px, py, pz, scale, 1.0f, r);
    if (first) {
        acc.v = w * r.v; acc.gx = w * r.gx; acc.gy = w * r.gy; acc.gz = w * r.gz;
    } else {
        acc.v = acc.v + w * r.v; acc.gx = acc.gx + w * r.gx; acc.gy = acc.gy + w * r.gy; acc.gz = acc.gz + w * r.gz;
    }
}
// Level l (0: the lower, 1: the upper) of the ladder at lr.
__device__ __forceinline__ Level ml_make(const FieldParams &P, const LevelRatio &lr, int l) {
    return make_level(P, l ? lr.shi : lr.slo, 1.0f, (uint32_t)P.level_seed_add[lr.li + 32 + l]);
}

// Normalised noise value and world gradient (evaluateNoise3D :142-165): world space, or
// with IRAY the isotropic ray space of the flight along (dx, dy, dz).
template <int NI, bool ISO, bool IRAY = false>
__device__ __forceinline__ Eval4 eval_noise(const FieldParams &P, const RealState &rs, 
                                            float px, float py, float pz, bool cond, float dx = 0.f, float dy = 0.f,
                                            float dz = 1.f) {
    if (IRAY) {
        // single-level SE: noise3D on frame.toLocal(p / l_aniso) with the identity kernel
        const Frame3 F = iso_ray_frame(P, dx, dy, dz);
        float rx, ry, rz;
        to_iso_ray(P, F, px, py, pz, rx, ry, rz);
        Eval4 n = noise3d_raw<NI, true>(P, rx, ry, rz, rs.seed, P.iso_R, P.iso_c0, P.iso_c0, P.iso_c0);
        float wx, wy, wz;
        from_iso_ray(P, F, n.gx * P.iso_g, n.gy * P.iso_g, n.gz * P.iso_g, wx, wy, wz);
        Eval4 r = {n.v * P.iso_nf_inv, wx * P.iso_nf_inv, wy * P.iso_nf_inv, wz * P.iso_nf_inv};
        if (cond && rs.cond) {
            // conditioning splats in the ray space: isotropic covariance kernel exp(-|d|^2 / 4)
            float ox, oy, oz;
            to_iso_ray(P, F, rs.ox, rs.oy, rs.oz, ox, oy, oz);
            const float ax = rx - ox, ay = ry - oy, az = rz - oz;
            const float k = __expf(-(ax * ax * 0.25f + ay * ay * 0.25f + az * az * 0.25f));
            const float d0 = -0.5f * ax, d1 = -0.5f * ay, d2 = -0.5f * az;
            const float bd = rs.bx * d0 + rs.by * d1 + rs.bz * d2;
            const float v = (rs.a + bd) * k;
            const float g0 = (rs.a * d0 + bd * d0 - 0.5f * rs.bx) * k;
            const float g1 = (rs.a * d1 + bd * d1 - 0.5f * rs.by) * k;
            const float g2 = (rs.a * d2 + bd * d2 - 0.5f * rs.bz) * k;
            float cx, cy, cz;
            from_iso_ray(P, F, g0, g1, g2, cx, cy, cz);
            r.v += v; r.gx += cx; r.gy += cy; r.gz += cz;
        }
        return r;
    }
    if (HasNoise<NI>::value && P.full) {
        // full anisotropy matrix: stationary SE, single level, world space
        Eval4 r = noise_full_normalized(P, px, py, pz, rs.seed);
        if (cond && rs.cond) cond_splat_full(P, rs, px, py, pz, r);
        return r;
    }
    if (!P.multires) {
        Eval4 r;
        float ls = 1.0f, gs = 1.0f;
        Level L;
        if (P.kind == 0) {
            L.R = P.R; L.c0 = P.c2[0]; L.c1 = P.c2[1]; L.c2 = P.c2[2];
            L.g0 = P.gfac[0]; L.g1 = P.gfac[1]; L.g2 = P.gfac[2]; L.nf_inv = P.nf_inv; L.seed_add = 0;
        } else {
            // brute-force nonstationary: gs = ls_max, ls = ls(p)/ls_max
            gs = P.ls_max;
            ls = ls_field<HasNoise<NI>::value>(P, px, py, pz) / P.ls_max;
            L = make_level(P, gs, ls, (uint32_t)P.seed_add_single);
            L.R = P.kernel_scale * gs * P.l_max;
        }
        Eval4 n = noise3d_raw<NI, ISO>(P, px, py, pz, rs.seed + L.seed_add, L.R, L.c0, L.c1, L.c2, true, rs.coop);
        r.v = n.v * L.nf_inv;
        r.gx = n.gx * L.g0 * L.nf_inv;
        r.gy = n.gy * L.g1 * L.nf_inv;
        r.gz = n.gz * L.g2 * L.nf_inv;
        if (cond && rs.cond) cond_splat(P, rs, px, py, pz, gs, ls, r);
        return r;
    }
    // multi-resolution: two adjacent levels of the base-2.5 ladder (kernelScaleLevelRatio)
    const LevelRatio lr = level_ratio(ls_field<HasNoise<NI>::value>(P, px, py, pz));
    Eval4 acc;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const Level L = ml_make(P, lr, l);
        const Eval4 n = noise3d_raw<NI, ISO>(P, px, py, pz, rs.seed + L.seed_add, L.R, L.c0, L.c1, L.c2, true, rs.coop);
        ml_level(P, rs, px, py, pz, cond, l ? lr.shi : lr.slo, l ? lr.whi : lr.wlo, L, n, l == 0, acc);
    }
    return acc;
}

// sparseConvNoiseAmplitude(p): sigma, or getVariance(p) * sigma for a nonstationary
// covariance (GPFunctions.cpp:1245-1247), whose "noise" variance field varies with p (NOISE
// kernels only, as ls_field).
template <bool NOISE>
__device__ __forceinline__ float amplitude(const FieldParams &P, float x, float y, float z) {
    if (NOISE && P.var_field == 1)
        return (float)(noise_field(P.var_noise, false, (double)x, (double)y, (double)z) * (double)P.sigma);
    return P.amp_const;
}

// f(p) = amplitude * noise + mean and its world gradient, rounded as evaluateValue and
// evaluateGradient round them (SparseConvolutionNoise.cpp:113-140): the float product
// amplitude * noise added to the double mean and rounded once; the gradient's float
// product and float mean gradient added unfused.  With an analytic (float) mean the double
// sum of two floats is exact, so the float addition rounds the same.
template <int MEAN, bool NOISE = false>
__device__ __forceinline__ void add_mean(const FieldParams &P, float px, float py, float pz, Eval4 &r) {
#pragma clang fp contract(off)
    float mgx, mgy, mgz;
    const float a = amplitude<NOISE>(P, px, py, pz);
    const float an = a * r.v;
    if (MEAN == MEAN_ANALYTIC) {
        r.v = an + analytic_mean_value(P, px, py, pz, mgx, mgy, mgz);
    } else {
        const double m = mean_value_d<MEAN>(P, px, py, pz, mgx, mgy, mgz);
        r.v = (float)((double)an + m);
    }
    r.gx = a * r.gx + mgx;
    r.gy = a * r.gy + mgy;
    r.gz = a * r.gz + mgz;
}

template <int NI, bool ISO, int MEAN = MEAN_ANY, bool IRAY = false>
__device__ __forceinline__ Eval4 eval_field(const FieldParams &P, const RealState &rs,
                                            float px, float py, float pz, bool cond, float dx = 0.f,
                                            float dy = 0.f, float dz = 1.f) {
    Eval4 r = eval_noise<NI, ISO, IRAY>(P, rs, px, py, pz, cond, dx, dy, dz);
    add_mean<MEAN, HasNoise<NI>::value>(P, px, py, pz, r);
    return r;
}

// f and its gradient at two points of one flight (a, and b when `two`): eval_field at each.
// Stationary single-level media in world space sum both points' noise in one pass over the
// support list of a (noise3d_raw2); the other media evaluate the points one after the other.
template <int NI, bool ISO, int MEAN = MEAN_ANY, bool IRAY = false>
__device__ __forceinline__ void eval_field2(const FieldParams &P, const RealState &rs, const float pa[3],
                                            const float pb[3], bool two, bool cond, float dx, float dy, float dz,
                                            Eval4 &A, Eval4 &B) {
    if (!IRAY && !P.multires && P.kind == 0 && !(HasNoise<NI>::value && P.full)) {
        Eval4 na, nb;
        noise3d_raw2<NI, ISO>(P, pa, pb, two, rs.seed, P.R, P.c2[0], P.c2[1], P.c2[2], na, nb);
        A.v = na.v * P.nf_inv;
        A.gx = na.gx * P.gfac[0] * P.nf_inv;
        A.gy = na.gy * P.gfac[1] * P.nf_inv;
        A.gz = na.gz * P.gfac[2] * P.nf_inv;
        B.v = nb.v * P.nf_inv;
        B.gx = nb.gx * P.gfac[0] * P.nf_inv;
        B.gy = nb.gy * P.gfac[1] * P.nf_inv;
        B.gz = nb.gz * P.gfac[2] * P.nf_inv;
#pragma unroll 1
        for (int w = 0; w < 2; ++w) {
            if (w == 1 && !two) break;
            Eval4 &E = w ? B : A;
            const float *p = w ? pb : pa;
            if (cond && rs.cond) cond_splat(P, rs, p[0], p[1], p[2], 1.0f, 1.0f, E);
            add_mean<MEAN, HasNoise<NI>::value>(P, p[0], p[1], p[2], E);
        }
        return;
    }
#pragma unroll 1
    for (int w = 0; w < 2; ++w) {
        if (w == 1 && !two) break;
        const float *p = w ? pb : pa;
        const Eval4 e = eval_field<NI, ISO, MEAN, IRAY>(P, rs, p[0], p[1], p[2], cond, dx, dy, dz);
        if (w) B = e;
        else A = e;
    }
}

}  // namespace gpis
