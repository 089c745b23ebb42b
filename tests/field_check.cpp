// Host build of the device field code (gpis_field.hpp, GPIS_HOST_CHECK: one lane per wave)
// with every table, list and fixed-point index checked against its array.  Built and run by
// tests/test_field_check.py.
//
// Checks, on a stationary SE field with the C1 length scale (R = 3 l / sqrt 2, l = 0.05,
// 8 impulses per cell, seed 1) over a table of the box [-1, 1]^3:
//   1. the impulse field of a query is the same, bit for bit, whether its cells come from
//      the generator (no table), a partial table (covering only x < 0) or the full table;
//   2. with support lists (1..6 sub-cells per cell edge) the support decisions are the
//      exact path's: with the splat set to 1 (exponent factor 0) the value is the integer
//      sum of the in-support weights, which must be identical, and the gradient sums agree;
//      with the SE exponent the value agrees to 2e-6 of the absolute sum;
//   3. every list holds each impulse within reach of its sub-cell exactly once, in the
//      reference's order, and nothing out of reach of the whole sub-cell;
//   4. the multi-resolution level index, level scales and weights of every float lateral
//      scale in [2^-12, 2^12] are those of the float logf/powf formula;
//   5. the 1D kernel's cells in LDS, made by PCG jump-ahead, equal the sequential
//      generator's after every kind of move along the ray;
//   6. the 3D cells made wave-cooperatively (one host thread per lane, 1..64 active lanes,
//      idle lanes among them; the shared rounds, and the per-lane loop of a wave with many
//      lanes holding points) give every lane the sequential loop's sum to the bit.
// Queries: uniform in the table's interior, near the origin planes (|p / R| < 1, where the
// fixed-point query is rounded), on cell and sub-cell faces, at distance 1 +- 2^-22 grid
// units from an impulse (the support radius), and up to four cells below the table's low
// corner (cells the table does not hold; their neighbour blocks must not be read from it).
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <random>
#include <thread>
#include <vector>

struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct uint4 { unsigned x, y, z, w; };

static long long g_checks = 0;
static void gpis_check_index(size_t i, size_t n, const char *file, int line) {
    ++g_checks;
    if (i >= n) {
        fprintf(stderr, "index %zu out of range %zu at %s:%d\n", i, n, file, line);
        abort();
    }
}

#define GPIS_HOST_CHECK 1
#define GPIS_LHD inline
#include "../gpis_extend_amd/csrc/gpis_field.hpp"

using namespace gpis;

struct Table {
    int lo[3], dim[3];
    std::vector<float> packed;
    std::vector<int32_t> fix;
    std::vector<uint32_t> ent, range;
    int S = 0;
    double reach = 1.0;
};

// The packed table of table_build_kernel: per cell N x (x with the weight's sign, y, z).
static void build_table(Table &T, const int lo[3], const int dim[3], uint32_t seed, int N) {
    for (int i = 0; i < 3; ++i) { T.lo[i] = lo[i]; T.dim[i] = dim[i]; }
    const long long cells = (long long)dim[0] * dim[1] * dim[2];
    T.packed.assign(cells * 3 * N, 0.f);
    for (long long c = 0; c < cells; ++c) {
        const int x = lo[0] + (int)(c % dim[0]), y = lo[1] + (int)((c / dim[0]) % dim[1]),
                  z = lo[2] + (int)(c / ((long long)dim[0] * dim[1]));
        Pcg g;
        g.set((uint64_t)(uint32_t)(xxh4((uint32_t)z, (uint32_t)y, (uint32_t)x, seed) + 1u));
        for (int k = 0; k < N; ++k) {
            const float iz = u01(g.next()), iy = u01(g.next()), ix = u01(g.next());
            const float w = u01(g.next()) < 0.5f ? -1.0f : 1.0f;
            T.packed[(c * N + k) * 3] = copysignf(ix, w);
            T.packed[(c * N + k) * 3 + 1] = iy;
            T.packed[(c * N + k) * 3 + 2] = iz;
        }
    }
}

// build_lists of gpis_hip.hip on the host: fixed-point table, counts, ranges, entries.
static void build_lists(Table &T, int N, int S, double reach) {
    T.S = S;
    T.reach = reach;
    const int m = lists::list_margin(reach);
    const long long cells = (long long)T.dim[0] * T.dim[1] * T.dim[2];
    const long long imps = cells * N;
    T.fix.assign((imps + 1) * 4, 0);
    for (long long i = 0; i < imps; ++i) {
        const long long c = i / N;
        const int x = T.lo[0] + (int)(c % T.dim[0]), y = T.lo[1] + (int)((c / T.dim[0]) % T.dim[1]),
                  z = T.lo[2] + (int)(c / ((long long)T.dim[0] * T.dim[1]));
        lists::fix_impulse(x, y, z, T.packed[3 * i], T.packed[3 * i + 1], T.packed[3 * i + 2], &T.fix[4 * i]);
    }
    lists::sentinel_impulse(&T.fix[4 * imps]);
    const long long nsub = cells * S * S * S;
    T.range.assign(nsub * 2, 0);
    unsigned long long total = 0;
    for (long long sc = 0; sc < nsub; ++sc) {
        const long long c = sc / (S * S * S);
        const int s = (int)(sc % (S * S * S));
        const int x = T.lo[0] + (int)(c % T.dim[0]), y = T.lo[1] + (int)((c / T.dim[0]) % T.dim[1]),
                  z = T.lo[2] + (int)(c / ((long long)T.dim[0] * T.dim[1]));
        uint32_t g = 0;
        if (lists::interior(T.lo, T.dim, x, y, z, m))
            g = (uint32_t)((lists::sub_cell_list(T.packed.data(), T.lo, T.dim, N, S, reach, x, y, z, s % S,
                                                 (s / S) % S, s / (S * S), nullptr, 0u) + 3) / 4);
        T.range[2 * sc] = (uint32_t)total;
        T.range[2 * sc + 1] = g;
        total += g;
    }
    // the lists, then one group of the sentinel alone (read for the groups past a list's end)
    T.ent.assign((total + 1) * 4, 0xFFFFFFFFu);
    for (int i = 0; i < 4; ++i) T.ent[total * 4 + i] = (uint32_t)(imps * 16);
    for (long long sc = 0; sc < nsub; ++sc) {
        if (!T.range[2 * sc + 1]) continue;
        const long long c = sc / (S * S * S);
        const int s = (int)(sc % (S * S * S));
        const int x = T.lo[0] + (int)(c % T.dim[0]), y = T.lo[1] + (int)((c / T.dim[0]) % T.dim[1]),
                  z = T.lo[2] + (int)(c / ((long long)T.dim[0] * T.dim[1]));
        const int n = lists::sub_cell_list(T.packed.data(), T.lo, T.dim, N, S, reach, x, y, z, s % S, (s / S) % S,
                                           s / (S * S), &T.ent[4 * (size_t)T.range[2 * sc]], (uint32_t)(imps * 16));
        assert((uint32_t)((n + 3) / 4) == T.range[2 * sc + 1]);
    }
    for (uint32_t e : T.ent) assert(e != 0xFFFFFFFFu && e % 16 == 0 && e <= imps * 16);
}

static FieldParams params(const Table *T, bool lists_on, float R, uint32_t seed, int N) {
    FieldParams P;
    memset(&P, 0, sizeof P);
    P.N = N;
    P.Nf = (float)N;
    P.R = R;
    if (T) {
        P.has_table = 1;
        P.table = T->packed.data();
        P.table_n = T->packed.size();
        P.tseed = seed;
        for (int i = 0; i < 3; ++i) { P.tlo[i] = T->lo[i]; P.tdim[i] = T->dim[i]; }
        if (lists_on && T->S) {
            P.has_lists = 1;
            P.lsub = T->S;
            P.lfix = T->fix.data();
            P.lent = T->ent.data();
            P.lrange = T->range.data();
            P.lfix_bytes = (uint32_t)(T->fix.size() * 4);
            P.lent_n = (uint32_t)T->ent.size();
            P.lrange_n = (uint32_t)(T->range.size() / 2);
            P.lmargin = lists::list_margin(T->reach);
            P.lpair2 = T->reach > 1.0 ? (float)((T->reach - 1.0) * (T->reach - 1.0) * 0.999) : 0.0f;
        }
    }
    return P;
}

static bool same_bits(const Eval4 &a, const Eval4 &b) {
    return memcmp(&a, &b, sizeof a) == 0;
}

int main(int argc, char **argv) {
    const int npts = argc > 1 ? atoi(argv[1]) : 20000;
    const int N = 8;
    const uint32_t seed = 1;
    const float l_conv = 0.05f * sqrtf(2.f) / 2;
    const float R = 3.0f * l_conv;
    const float c_se = -1.4426950408889634f * (R * R * (1.0f / (l_conv * l_conv) * 0.5f));
    int lo[3], dim[3];
    for (int i = 0; i < 3; ++i) {
        lo[i] = (int)floorf(-1.0f / R) - 1;
        dim[i] = (int)floorf(1.0f / R) + 1 - lo[i] + 1;
    }
    Table full;
    build_table(full, lo, dim, seed, N);
    // the partial table: cells with x < 0 only
    Table part;
    {
        int plo[3] = {lo[0], lo[1], lo[2]}, pdim[3] = {-lo[0], dim[1], dim[2]};
        build_table(part, plo, pdim, seed, N);
    }

    std::mt19937 rng(2024);
    std::uniform_real_distribution<float> U(-0.999f, 0.999f), Usmall(-0.99f * R, 0.99f * R), U01(0.f, 1.f);
    std::vector<float> pts;
    for (int i = 0; i < npts; ++i) {
        float p[3];
        const int kind = i % 5;
        for (int a = 0; a < 3; ++a) p[a] = U(rng);
        if (kind == 1) p[i % 3] = Usmall(rng);                       // near an origin plane
        if (kind == 4 && (i / 5) % 4 == 0) p[i % 3] = -1.0f - 0.45f * U01(rng);  // below the table's low corner
        if (kind == 2) {                                             // on a cell or sub-cell face
            const int a = i % 3;
            const float g = floorf(p[a] / R) + (float)(rng() % 7) / 6.0f;
            p[a] = g * R;
        }
        if (kind == 3) {                                             // at the support radius of an impulse
            const long long c = rng() % (full.packed.size() / (3 * N));
            const int k = rng() % N;
            const int cx = lo[0] + (int)(c % dim[0]), cy = lo[1] + (int)((c / dim[0]) % dim[1]),
                      cz = lo[2] + (int)(c / ((long long)dim[0] * dim[1]));
            const double ix = cx + fabs(full.packed[(c * N + k) * 3]), iy = cy + full.packed[(c * N + k) * 3 + 1],
                         iz = cz + full.packed[(c * N + k) * 3 + 2];
            double ux = U01(rng) - 0.5, uy = U01(rng) - 0.5, uz = U01(rng) - 0.5;
            const double ul = sqrt(ux * ux + uy * uy + uz * uz);
            const double rad = 1.0 + ((int)(rng() % 5) - 2) * ldexp(1.0, -22);
            p[0] = (float)((ix + rad * ux / ul) * R);
            p[1] = (float)((iy + rad * uy / ul) * R);
            p[2] = (float)((iz + rad * uz / ul) * R);
            if (fabsf(p[0]) > 0.999f || fabsf(p[1]) > 0.999f || fabsf(p[2]) > 0.999f) { --i; continue; }
        }
        pts.insert(pts.end(), p, p + 3);
    }

    // 1. generator, partial table and full table agree bit for bit
    const FieldParams Pn = params(nullptr, false, R, seed, N), Pp = params(&part, false, R, seed, N),
                      Pf = params(&full, false, R, seed, N);
    long long table_mismatch = 0;
    for (int i = 0; i < npts; ++i) {
        const float *p = &pts[3 * i];
        const Eval4 a = noise3d_raw<8, true>(Pn, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
        const Eval4 b = noise3d_raw<8, true>(Pp, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
        const Eval4 c = noise3d_raw<8, true>(Pf, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
        const Eval4 d = noise3d_raw<0, true>(Pf, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
        if (!same_bits(a, b) || !same_bits(a, c) || !same_bits(a, d)) ++table_mismatch;
    }
    printf("table_mismatch %lld\n", table_mismatch);

    // 2./3. lists
    long long support_mismatch = 0, value_bad = 0, near = 0, evals = 0, near_radius = 0, evals_radius = 0;
    double max_rel = 0.0, mean_list = 0.0;
    long long list_cnt = 0, list_missing = 0, list_extra = 0;
    long long pair_mismatch = 0, pair_evals = 0;
    // sub-cells per edge 1, 3, 6 and 8; reach one grid unit, the pair reach of this field's
    // step (0.01: 1.095 grid units) and the largest pair reach
    const int subs[4] = {1, 3, 6, 8};
    const double reaches[3] = {1.0, lists::pair_reach(0.01, (double)R), (double)lists::kPairReach};
    for (int variant = 0; variant < 12; ++variant) {
        const int S = subs[variant % 4];
        const double reach = reaches[variant / 4];
        const int mg = lists::list_margin(reach);
        build_lists(full, N, S, reach);
        const FieldParams Pl = params(&full, true, R, seed, N);
        for (int i = 0; i < npts; ++i) {
            const float *p = &pts[3 * i];
            // indicator splats: v = sum of in-support weights (exact integers on both paths)
            const Eval4 e0 = noise3d_raw<8, true>(Pf, p[0], p[1], p[2], seed, R, 0.f, 0.f, 0.f);
            const Eval4 e1 = noise3d_raw<8, true>(Pl, p[0], p[1], p[2], seed, R, 0.f, 0.f, 0.f);
            if (e0.v != e1.v || fabsf(e0.gx - e1.gx) > 1e-5f * (1.f + fabsf(e0.gx)) ||
                fabsf(e0.gy - e1.gy) > 1e-5f * (1.f + fabsf(e0.gy)) || fabsf(e0.gz - e1.gz) > 1e-5f * (1.f + fabsf(e0.gz)))
                ++support_mismatch;
            const Eval4 s0 = noise3d_raw<8, true>(Pf, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
            const Eval4 s1 = noise3d_raw<8, true>(Pl, p[0], p[1], p[2], seed, R, c_se, c_se, c_se);
            const double rel = fabs((double)s0.v - s1.v) + fabs((double)s0.gx - s1.gx) +
                               fabs((double)s0.gy - s1.gy) + fabs((double)s0.gz - s1.gz);
            if (rel > 2e-5) ++value_bad;
            if (rel > max_rel) max_rel = rel;
            // near flags of the list path (the lane would run the exact path)
            float gx = p[0] / R, gy = p[1] / R, gz = p[2] / R;
            float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
            int ix = (int)fx, iy = (int)fy, iz = (int)fz;
            if (lists::interior(full.lo, full.dim, ix, iy, iz, mg)) {
                const unsigned lc = (unsigned)lists::cell_linear(full.lo, full.dim, ix, iy, iz);
                float v, ax, ay, az;
                const bool nr =
                    list_sum<true>(Pl, lc, gx, gy, gz, gx - fx, gy - fy, gz - fz, c_se, c_se, c_se, v, ax, ay, az);
                if (i % 5 == 3) {
                    near_radius += nr;  // queries placed at the support radius
                    ++evals_radius;
                } else {
                    near += nr;
                    ++evals;
                }
            }
        }
        // two queries per pass: b within the lists' extra reach of a, against each on its own
        if (Pl.lpair2 > 0.0f) {
            for (int i = 0; i < npts; ++i) {
                const float *p = &pts[3 * i];
                float u[3] = {U01(rng) - 0.5f, U01(rng) - 0.5f, U01(rng) - 0.5f};
                const float ul = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                const float len = sqrtf(Pl.lpair2) * U01(rng) * R;
                const float pb[3] = {p[0] + len * u[0] / ul, p[1] + len * u[1] / ul, p[2] + len * u[2] / ul};
                for (int cz = 0; cz < 2; ++cz) {
                    const float cc = cz ? c_se : 0.f;
                    Eval4 A, B;
                    noise3d_raw2<8, true>(Pl, p, pb, true, seed, R, cc, cc, cc, A, B);
                    const Eval4 a0 = noise3d_raw<8, true>(Pf, p[0], p[1], p[2], seed, R, cc, cc, cc);
                    const Eval4 b0 = noise3d_raw<8, true>(Pf, pb[0], pb[1], pb[2], seed, R, cc, cc, cc);
                    const double d = fabs((double)A.v - a0.v) + fabs((double)B.v - b0.v) + fabs((double)A.gx - a0.gx) +
                                     fabs((double)B.gz - b0.gz);
                    if ((cz == 0 && (A.v != a0.v || B.v != b0.v)) || d > 2e-5) ++pair_mismatch;
                    ++pair_evals;
                    // a query without a second point
                    Eval4 A1, B1;
                    noise3d_raw2<8, true>(Pl, p, pb, false, seed, R, cc, cc, cc, A1, B1);
                    if ((cz == 0 && A1.v != a0.v) || fabs((double)A1.v - a0.v) + fabs((double)A1.gy - a0.gy) > 2e-5)
                        ++pair_mismatch;
                }
            }
        }
        // list contents against a brute-force scan of every interior cell's sub-cells
        const int step = S >= 4 ? 7 : 1;
        mean_list = 0.0;
        list_cnt = 0;
        for (long long c = 0; c < (long long)dim[0] * dim[1] * dim[2]; c += step) {
            const int x = lo[0] + (int)(c % dim[0]), y = lo[1] + (int)((c / dim[0]) % dim[1]),
                      z = lo[2] + (int)(c / ((long long)dim[0] * dim[1]));
            if (!lists::interior(lo, dim, x, y, z, mg)) continue;
            for (int s = 0; s < S * S * S; ++s) {
                const long long sc = c * S * S * S + s;
                const uint32_t *e = &full.ent[4 * (size_t)full.range[2 * sc]];
                const int ng = (int)full.range[2 * sc + 1];
                const int sx = s % S, sy = (s / S) % S, sz = s / (S * S);
                int at = 0;
                for (int dx = -mg; dx <= mg; ++dx)
                    for (int dy = -mg; dy <= mg; ++dy)
                        for (int dz = -mg; dz <= mg; ++dz)
                            for (int k = 0; k < N; ++k) {
                                const long long cn = lists::cell_linear(lo, dim, x + dx, y + dy, z + dz);
                                const float *q = &full.packed[(cn * N + k) * 3];
                                // exact distance from the impulse to the unpadded sub-cell box
                                double d2 = 0;
                                const double e3[3] = {dx + (double)fabsf(q[0]), dy + (double)q[1], dz + (double)q[2]};
                                const int s3[3] = {sx, sy, sz};
                                for (int a = 0; a < 3; ++a) {
                                    const double b0 = (double)s3[a] / S, b1 = (double)(s3[a] + 1) / S;
                                    const double g = e3[a] < b0 ? b0 - e3[a] : (e3[a] > b1 ? e3[a] - b1 : 0.0);
                                    d2 += g * g;
                                }
                                const uint32_t off = (uint32_t)((cn * N + k) * 16);
                                const bool listed = at < 4 * ng && e[at] == off;
                                if (listed) ++at;
                                if (d2 < reach * reach && !listed) ++list_missing;
                                if (d2 > reach * reach * 1.001 && listed) ++list_extra;
                            }
                for (; at < 4 * ng; ++at)
                    if (e[at] != (uint32_t)(full.packed.size() / 3 * 16)) ++list_extra;  // only sentinels remain
                list_cnt += 1;
                mean_list += 4.0 * ng;
            }
        }
        printf("reach %.3f S %d support_mismatch %lld value_bad %lld near %lld of %lld (at the radius: %lld of %lld) "
               "max_abs_diff %.3g mean_list %.1f\n", reach, S, support_mismatch, value_bad, near, evals, near_radius,
               evals_radius, max_rel, mean_list / (list_cnt ? list_cnt : 1));
    }
    printf("pair_mismatch %lld of %lld\n", pair_mismatch, pair_evals);
    printf("list_missing %lld list_extra %lld checks %lld\n", list_missing, list_extra, g_checks);

    // 4. multi-resolution levels: float lateral scales in [2^-12, 2^12] (every `lstep`-th
    //    float, and every float within 2^14 of each level boundary 2.5^L) get the level
    //    index and the scales 2.5^L, 2.5^(L+1) of the float logf/powf formula
    //    (SparseConvolutionNoise.cpp:101-111), bit for bit, and its weights
    long long level_bad = 0, weight_bad = 0, level_n = 0;
    {
        auto check_level = [&](uint32_t u) {
            float lat;
            memcpy(&lat, &u, 4);
            const LevelRatio k = level_ratio(lat);
            const float lo = floorf(logf(lat) / logf(2.5f)), slo = powf(2.5f, lo), shi = powf(2.5f, lo + 1.0f);
            const float fr = (lat - slo) / (shi - slo);
            const float wlo = (float)((1.0 - fr) / sqrt(1.0 - 2.0 * fr + 2.0 * fr * fr));
            const float whi = (float)(fr / sqrt(1.0 - 2.0 * fr + 2.0 * fr * fr));
            if (k.li != (int)lo || k.slo != slo || k.shi != shi) {
                if (level_bad < 4) printf("level: lat %.9g device (%d, %.9g, %.9g) float formula (%d, %.9g, %.9g)\n",
                                          lat, k.li, k.slo, k.shi, (int)lo, slo, shi);
                ++level_bad;
            }
            if (k.wlo != wlo || k.whi != whi) ++weight_bad;
            ++level_n;
        };
        const int lstep = argc > 2 ? atoi(argv[2]) : 1;
        uint32_t u0, u1;
        const float f0 = ldexpf(1.0f, -12), f1 = ldexpf(1.0f, 12);
        memcpy(&u0, &f0, 4);
        memcpy(&u1, &f1, 4);
        for (uint32_t u = u0; u <= u1; u += (uint32_t)lstep) check_level(u);
        for (int L = -9; L <= 9; ++L) {
            const float b = powf(2.5f, (float)L);
            uint32_t ub;
            memcpy(&ub, &b, 4);
            for (uint32_t u = ub - (1u << 14); u <= ub + (1u << 14); ++u) check_level(u);
        }
    }
    printf("level_mismatch %lld weight_mismatch %lld of %lld\n", level_bad, weight_bad, level_n);

    // 5. the 1D flight kernel's cells in LDS, made through PCG jump-ahead: after every move of
    // a lane (steps of 0, +-1, +-2 cells, jumps, a new flight) its ring slots hold cells
    // i-1, i, i+1 exactly as the sequential generator makes them
    long long cells1d_bad = 0, cells1d_n = 0;
    {
        static float4 lds[3 * 2 * 64];
        static uint32_t jobs[2 * Cells1DShared<8>::kMaxJobs];
        Cells1DShared<8> C;
        C.cells = lds; C.job_st = jobs; C.job_dst = jobs + Cells1DShared<8>::kMaxJobs;
        C.base = 0; C.valid = false;
        std::mt19937 r1(7);
        uint32_t seed1 = 0x9E3779B9u;
        int i = -40;
        for (int it = 0; it < 20000; ++it) {
            const int pick = (int)(r1() % 16);
            if (pick < 5) i += 0;
            else if (pick < 9) i += 1;
            else if (pick < 11) i -= 1;
            else if (pick < 12) i += 2;
            else if (pick < 13) i -= 2;
            else if (pick < 15) i += (int)(r1() % 97) - 48;
            else { C.valid = false; seed1 = (uint32_t)r1(); }
            cells_shared_for<8>(C, i, seed1);
            for (int dx = 0; dx < 3; ++dx) {
                float ref[8];
                gen_cell1d<8>(ref, i - 1 + dx, seed1);
                const int slot = ring_slot(i - 1 + dx);
                for (int k = 0; k < 8; ++k) {
                    const float4 q = lds[(slot * 2 + k / 4) * 64 + 0];
                    const float got = (&q.x)[k & 3];
                    if (memcmp(&got, &ref[k], 4) != 0) ++cells1d_bad;
                    ++cells1d_n;
                }
            }
        }
    }
    printf("cells1d_mismatch %lld of %lld\n", cells1d_bad, cells1d_n);
    // 6. wave-cooperative generated cells (cells_coop): one thread per lane of a wave of 1..64
    //    active lanes (a random subset of the 64 lane ids), some of them idle; every lane's
    //    sum equals the sequential loop's to the bit, an idle lane's is zero
    long long coop_bad = 0, coop_n = 0;
    {
        std::mt19937 rc(11);
        std::vector<float4> lds(coop_wave_slots(kCoopMaxN));
        for (int trial = 0; trial < 180; ++trial) {
            const int pick = trial % 6;
            const int nact = pick == 0 ? 64 : pick == 1 ? 1 : pick == 2 ? 2 : 1 + (int)(rc() % 64);
            const int n_imp = trial % 2 ? 8 : 10;  // NI = 8, and the generic count with 10
            const float idle_p = (float)(trial % 4) / 4.0f;
            std::vector<unsigned> ids(64);
            for (unsigned i = 0; i < 64; ++i) ids[i] = i;
            std::shuffle(ids.begin(), ids.end(), rc);
            ids.resize(nact);
            std::sort(ids.begin(), ids.end());
            struct LaneIn { float p[3], R; uint32_t seed; bool idle; Eval4 got, ref; };
            std::vector<LaneIn> in(nact);
            std::uniform_real_distribution<float> u(-1.0f, 1.0f), ur(0.03f, 0.2f), u01f(0.0f, 1.0f);
            for (auto &L : in) {
                for (float &c : L.p) c = u(rc);
                L.R = ur(rc);
                L.seed = (uint32_t)rc();
                L.idle = u01f(rc) < idle_p;
            }
            FieldParams Pc = params(nullptr, false, 0.1f, 1u, n_imp);
            host_coop_max_live = trial % 3 == 0 ? kCoopMaxLive : 64;  // the per-lane loop, or always shared
            HostWave W;
            pthread_barrier_init(&W.bar, nullptr, (unsigned)nact);
            W.vote = 0;
            W.first = ids[0];
            host_wave = &W;
            host_coop_lds = lds.data();
            std::vector<std::thread> th;
            for (int l = 0; l < nact; ++l) {
                th.emplace_back([&, l] {
                    host_lane = ids[l];
                    LaneIn &L = in[l];
                    const float c = -1.4426950408889634f * 2.25f;
                    if (n_imp == 8)
                        L.got = noise3d_raw<8, true>(Pc, L.p[0], L.p[1], L.p[2], L.seed, L.R, c, c, c, true, L.idle ? 2 : 1);
                    else
                        L.got = noise3d_raw<0, true>(Pc, L.p[0], L.p[1], L.p[2], L.seed, L.R, c, c, c, true, L.idle ? 2 : 1);
                });
            }
            for (auto &t : th) t.join();
            host_wave = nullptr;
            host_coop_lds = nullptr;
            pthread_barrier_destroy(&W.bar);
            for (auto &L : in) {
                const float c = -1.4426950408889634f * 2.25f;
                const Eval4 zero = {0.f, 0.f, 0.f, 0.f};
                if (L.idle) L.ref = zero;
                else if (n_imp == 8) L.ref = noise3d_raw<8, true>(Pc, L.p[0], L.p[1], L.p[2], L.seed, L.R, c, c, c);
                else L.ref = noise3d_raw<0, true>(Pc, L.p[0], L.p[1], L.p[2], L.seed, L.R, c, c, c);
                if (!same_bits(L.got, L.ref)) {
                    if (coop_bad < 4)
                        printf("coop: trial %d lanes %d lane value %.9g ref %.9g\n", trial, nact, L.got.v, L.ref.v);
                    ++coop_bad;
                }
                ++coop_n;
            }
        }
    }
    printf("coop_mismatch %lld of %lld\n", coop_bad, coop_n);
    const bool ok = level_bad == 0 && weight_bad == 0 && table_mismatch == 0 && support_mismatch == 0 && pair_mismatch == 0 && pair_evals > 0 && value_bad == 0 && list_missing == 0 &&
                    list_extra == 0 && near * 100 < evals && near_radius * 10 > evals_radius * 9 && cells1d_bad == 0 &&
                    cells1d_n > 0 && coop_bad == 0 && coop_n > 0;
    printf("%s\n", ok ? "OK" : "FAIL");
    return ok ? 0 : 1;
}
