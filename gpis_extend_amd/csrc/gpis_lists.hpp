// gpis_lists.hpp — support lists of a single-realization impulse table.
//
// cell3D (SparseConvolutionNoise.cpp:432-452) keeps an impulse only when the query lies
// within one grid unit of it (|to|^2 < 1), and noise3D (:419-430) visits the 27 cells
// around the query.  Of their 27N impulses, a query has on average (4 pi / 3) N in
// support.  A lane of a wave runs a cell whenever any lane needs it, so with lanes at
// unrelated points every evaluation costs all 27N support tests.
//
// A support list names, for one sub-cell (the cell cut S times along each axis), the
// impulses of the 27 neighbour cells that can be in support for some query inside the
// sub-cell: those whose distance to the sub-cell box is below one grid unit (with a margin
// far above rounding).  The impulses outside the list are out of support for every query
// of the sub-cell, so a lane that sums over its own list only skips exact zeros.  Lists
// keep the reference's visiting order (dx, then dy, then dz, then the impulse), so the sum
// adds the same terms in the same order.  At S = 3 a list holds about 64 impulses of 216.
//
// Each list entry is a byte offset into a fixed-point copy of the table: per impulse the
// absolute grid coordinates times 2^23 as int32 (cell index * 2^23 + the 23-bit mantissa
// of normalizedUint, BitManip.hpp:47-50), and the weight as a float.  The difference of a
// query's fixed-point coordinate and an impulse's is exact in int32, and its conversion to
// float is the correctly rounded difference: for a query whose grid coordinate has no bits
// below 2^-23 (|p / R| >= 1) this is the reference's to = (frac - d) - p_i to the bit,
// whatever the neighbour offset d.  Near the origin planes (|p / R| < 1) the query is
// rounded to 2^-23 first; the lane then flags any pair whose |to|^2 lies within a relative
// 2e-6 of 1 and the caller decides that query again on the exact 27-cell path, so the
// support decisions stay the reference's for every query.
//
// Lists are padded to groups of four entries with a sentinel impulse placed 2^7 cells away
// (out of support for every query).  Per sub-cell: the index of its first group and the
// number of groups.
#pragma once

#include <math.h>
#include <stdint.h>

#ifndef GPIS_LHD
#define GPIS_LHD __host__ __device__ __forceinline__
#endif

namespace gpis {
namespace lists {

constexpr float kFixScale = 8388608.0f;      // 2^23
constexpr int kMaxAbsCell = 120;             // |cell index| bound: query - impulse and query - sentinel fit int32
constexpr int32_t kSentinel = 1 << 30;       // sentinel impulse coordinate (2^7 cells away)
constexpr double kReach2 = 1.0001;           // list membership: squared distance margin (relative)
constexpr float kPairReach = 1.125f;         // largest reach of lists that serve two queries per pass
constexpr double kBoxPad = 1e-5;             // sub-cell box padding (rounding of frac * S)
// Pairs with | |to|^2 - 1 | below kNearRel are decided again on the exact path.  The list
// path's |to|^2 differs from the reference's by the fused products (a few ulp, < 3e-7) and,
// for queries with |p / R| < 1, by the rounding of the query to 2^-23 (|to| moves by at most
// 2^-23 per axis: < 8e-7); 2e-6 covers both with a margin.
constexpr float kNearRel = 2e-6f;
constexpr int kMaxSub = 8;                   // sub-cells per cell edge: 1..kMaxSub

// Reach of lists that serve two queries per pass: one grid unit plus the farthest second
// query of a flight, the march's lattice step (step_size / R grid units for a unit
// direction) with a margin, at most kPairReach.  A pair farther apart than the lists allow
// is summed as two passes, so the reach only sets how many pairs share a pass.
GPIS_LHD double pair_reach(double step_size, double R) {
    const double e = step_size / R * 1.01 + 1e-4;
    return 1.0 + (e < (double)kPairReach - 1.0 ? e : (double)kPairReach - 1.0);
}

// Linear index of the table cell (x, y, z) (absolute cell indices), x fastest.
GPIS_LHD long long cell_linear(const int lo[3], const int dim[3], int x, int y, int z) {
    return ((long long)(z - lo[2]) * dim[1] + (y - lo[1])) * dim[0] + (x - lo[0]);
}

// Interior cells: every cell within m cells along each axis lies in the table.
GPIS_LHD bool interior(const int lo[3], const int dim[3], int x, int y, int z, int m) {
    return x - m >= lo[0] && x + m < lo[0] + dim[0] && y - m >= lo[1] && y + m < lo[1] + dim[1] &&
           z - m >= lo[2] && z + m < lo[2] + dim[2];
}
// Neighbourhood of a list: one cell around the query's cell for a reach of one grid unit,
// two when the list also serves a second query up to (reach - 1) away.
GPIS_LHD int list_margin(double reach) { return reach > 1.0 ? 2 : 1; }

// Fixed-point impulse k of cell (cx, cy, cz) from its packed form (x with the weight in its
// sign bit, y, z): out = (X, Y, Z, bits of the weight).
GPIS_LHD void fix_impulse(int cx, int cy, int cz, float xs, float y, float z, int32_t out[4]) {
    const float ax = fabsf(xs);
    out[0] = cx * (1 << 23) + (int32_t)(ax * kFixScale);
    out[1] = cy * (1 << 23) + (int32_t)(y * kFixScale);
    out[2] = cz * (1 << 23) + (int32_t)(z * kFixScale);
    const float w = xs < 0.0f || (xs == 0.0f && signbit(xs)) ? -1.0f : 1.0f;
    union { float f; int32_t i; } u;
    u.f = w;
    out[3] = u.i;
}
GPIS_LHD void sentinel_impulse(int32_t out[4]) {
    out[0] = out[1] = out[2] = kSentinel;
    out[3] = 0;
}

// Squared distance from the point e to the box [lo, hi] along one axis.
GPIS_LHD double gap2(double e, double lo, double hi) {
    const double g = e < lo ? lo - e : (e > hi ? e - hi : 0.0);
    return g * g;
}

// The support list of sub-cell (sx, sy, sz) of interior cell (cx, cy, cz): the number of
// entries (before padding), and when `out` is given the entries themselves (byte offsets of
// fixed-point impulses; impulse i of the table is at 16 i) padded with `sentinel_off` to a
// multiple of four.  `table` is the packed table (3N floats per cell).  An impulse is listed
// when its distance to the sub-cell box is below `reach` grid units (1: the support of the
// queries inside the sub-cell; 1 + e: also of every query within e of one inside it).  The
// cells are visited in the order of noise3D's loops (x, then y, then z), which orders the
// list the same way for every query it serves.
GPIS_LHD int sub_cell_list(const float *table, const int lo[3], const int dim[3], int N, int S, double reach, int cx,
                           int cy, int cz, int sx, int sy, int sz, uint32_t *out, uint32_t sentinel_off) {
    const double bx0 = (double)sx / S - kBoxPad, bx1 = (double)(sx + 1) / S + kBoxPad;
    const double by0 = (double)sy / S - kBoxPad, by1 = (double)(sy + 1) / S + kBoxPad;
    const double bz0 = (double)sz / S - kBoxPad, bz1 = (double)(sz + 1) / S + kBoxPad;
    const double r2 = reach * reach * kReach2;
    const int m = list_margin(reach);
    int n = 0;
    for (int dx = -m; dx <= m; ++dx)
        for (int dy = -m; dy <= m; ++dy)
            for (int dz = -m; dz <= m; ++dz) {
                // cells whose box is out of reach of the sub-cell box hold no listed impulse
                const double cb = gap2(dx + 0.5, bx0 - 0.5, bx1 + 0.5) + gap2(dy + 0.5, by0 - 0.5, by1 + 0.5) +
                                  gap2(dz + 0.5, bz0 - 0.5, bz1 + 0.5);
                if (cb >= r2) continue;
                const long long c = cell_linear(lo, dim, cx + dx, cy + dy, cz + dz);
                const float *q = table + c * 3 * N;
                for (int k = 0; k < N; ++k) {
                    const double ex = dx + (double)fabsf(q[3 * k]);
                    const double ey = dy + (double)q[3 * k + 1];
                    const double ez = dz + (double)q[3 * k + 2];
                    const double d2 = gap2(ex, bx0, bx1) + gap2(ey, by0, by1) + gap2(ez, bz0, bz1);
                    if (d2 < r2) {
                        if (out) out[n] = (uint32_t)((c * N + k) * 16);
                        ++n;
                    }
                }
            }
    if (out)
        for (int i = n; i % 4 != 0; ++i) out[i] = sentinel_off;
    return n;
}

// Sub-cell of a query with in-cell coordinate r in [0, 1).
GPIS_LHD int sub_index(float r, int S) {
    const int s = (int)(r * (float)S);
    return s < 0 ? 0 : (s > S - 1 ? S - 1 : s);
}

}  // namespace lists
}  // namespace gpis
