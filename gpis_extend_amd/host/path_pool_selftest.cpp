// path_pool_selftest.cpp — runs PathPool over HipRayBatchBackend for a file of path
// requests and writes the per-path results for the caller to compare with the
// device-resident paths of gpis_trace_paths_1d.
//
//   path_pool_selftest <desc.bin> <paths.bin> <n> <results.bin> <max_bounces> <phase>
//                      <bounce_seed> <box> <scene_seed> <pool> <batch> [lx ly lz]
//
// paths.bin: n records of 20 four-byte words: origin[3], direction[3], nearT, farT, jitter,
// flags (bit 0 firstScatter), previousValue, previousGradient[3], pixelSampleSegment[4],
// 4 unused.  results.bin: n records of 12 words: t0, gradient0[3], outcome0, flights, hits,
// shadows, visible, nee, tLast, outcomeLast.  stdout: rounds, batches, segments.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "path_pool.hpp"

using namespace gpis_host;

int main(int argc, char **argv) {
    // an optional last argument "exact" or "fast" selects the 1D arithmetic
    int precision = GPIS_PRECISION_FAST;
    if (argc > 1 && (std::strcmp(argv[argc - 1], "exact") == 0 || std::strcmp(argv[argc - 1], "fast") == 0)) {
        precision = std::strcmp(argv[argc - 1], "exact") == 0 ? GPIS_PRECISION_EXACT : GPIS_PRECISION_FAST;
        --argc;
    }
    if (argc < 12) {
        std::fprintf(stderr, "usage: %s desc.bin paths.bin n results.bin max_bounces phase bounce_seed box "
                             "scene_seed pool batch [lx ly lz]\n", argv[0]);
        return 2;
    }
    gpis_field_desc desc;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(&desc, sizeof desc, 1, f) != 1) { std::fprintf(stderr, "bad desc\n"); return 2; }
    std::fclose(f);
    const long n = std::atol(argv[3]);
    std::vector<std::uint32_t> raw((size_t)n * 20);
    f = std::fopen(argv[2], "rb");
    if (!f || std::fread(raw.data(), 4, raw.size(), f) != raw.size()) { std::fprintf(stderr, "bad paths\n"); return 2; }
    std::fclose(f);
    std::vector<PathRequest> paths((size_t)n);
    for (long i = 0; i < n; ++i) {
        const std::uint32_t *w = &raw[(size_t)i * 20];
        float v[20];
        std::memcpy(v, w, sizeof v);
        PathRequest &p = paths[(size_t)i];
        for (int a = 0; a < 3; ++a) { p.origin[a] = v[a]; p.direction[a] = v[3 + a]; p.previousGradient[a] = v[11 + a]; }
        p.nearT = v[6]; p.farT = v[7]; p.jitter = v[8];
        p.firstScatter = (w[9] & 1u) != 0;
        p.previousValue = v[10];
        for (int a = 0; a < 4; ++a) p.pixelSampleSegment[a] = w[14 + a];
    }
    PathPoolSettings s;
    s.maxBounces = std::atoi(argv[5]);
    s.phase = std::atoi(argv[6]);
    s.bounceSeed = (std::uint32_t)std::strtoul(argv[7], nullptr, 10);
    s.box = (float)std::atof(argv[8]);
    s.sceneSeed = (std::uint32_t)std::strtoul(argv[9], nullptr, 10);
    s.poolSize = (std::size_t)std::atol(argv[10]);
    s.batchSize = (std::size_t)std::atol(argv[11]);
    if (argc >= 15) {
        s.useLight = true;
        for (int a = 0; a < 3; ++a) s.light[a] = (float)std::atof(argv[12 + a]);
    }
    HipRayBatchBackend be(desc, s.batchSize);
    if (!be.ok() || !be.setPrecision(precision)) { std::fprintf(stderr, "backend: %s\n", be.error().c_str()); return 3; }
    PathPool pool(be, desc, s);
    std::vector<PathResult> res;
    if (!pool.run(paths, res)) { std::fprintf(stderr, "path pool: %s\n", pool.error().c_str()); return 4; }
    std::vector<std::uint32_t> out((size_t)n * 12);
    for (long i = 0; i < n; ++i) {
        const PathResult &r = res[(size_t)i];
        std::uint32_t *o = &out[(size_t)i * 12];
        std::memcpy(&o[0], &r.t0, 4);
        std::memcpy(&o[1], r.gradient0, 12);
        o[4] = r.outcome0; o[5] = r.flights; o[6] = r.hits; o[7] = r.shadows; o[8] = r.visible;
        std::memcpy(&o[9], &r.nee, 4);
        std::memcpy(&o[10], &r.tLast, 4);
        o[11] = r.outcomeLast;
    }
    f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(out.data(), 4, out.size(), f) != out.size()) { std::fprintf(stderr, "write failed\n"); return 2; }
    std::fclose(f);
    std::printf("rounds %llu batches %llu segments %llu\n", (unsigned long long)pool.stats().rounds,
                (unsigned long long)pool.stats().batches, (unsigned long long)pool.stats().segments);
    return 0;
}
