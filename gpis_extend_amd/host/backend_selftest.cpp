// backend_selftest.cpp — drives HipRayBatchBackend the way a path-pool coordinator
// would: several submitBatch calls in flight, replies awaited in reverse order, an
// invalid request mixed in, onComplete counted.  Reads requests from a binary file
// (16 x 4-byte fields per request, SoA by request) and writes the replies
// (t, value, gx, gy, gz, outcome, evaluations, valid) for the caller to compare.
//
//   backend_selftest <desc.bin> <requests.bin> <n> <replies.bin> [lo hi seed]
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpis_backend.hpp"

using namespace gpis_host;

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s desc.bin requests.bin n replies.bin [lo hi seed]\n", argv[0]);
        return 2;
    }
    gpis_field_desc desc;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(&desc, sizeof desc, 1, f) != 1) { std::fprintf(stderr, "bad desc\n"); return 2; }
    std::fclose(f);
    const long n = std::atol(argv[3]);
    std::vector<float> raw((size_t)n * 16);
    f = std::fopen(argv[2], "rb");
    if (!f || std::fread(raw.data(), 4, raw.size(), f) != raw.size()) { std::fprintf(stderr, "bad requests\n"); return 2; }
    std::fclose(f);
    std::vector<RaySegmentRequest> reqs((size_t)n);
    for (long i = 0; i < n; ++i) {
        const float *r = &raw[(size_t)i * 16];
        RaySegmentRequest &q = reqs[(size_t)i];
        for (int a = 0; a < 3; ++a) { q.origin[a] = r[a]; q.direction[a] = r[3 + a]; q.previousGradient[a] = r[13 + a]; }
        q.nearT = r[6]; q.farT = r[7]; q.jitter = r[8];
        std::memcpy(&q.noiseSeed, &r[9], 4);
        std::memcpy(&q.normalSeed, &r[10], 4);
        std::uint32_t flags;
        std::memcpy(&flags, &r[11], 4);
        q.firstScatter = (flags & 1u) != 0;
        q.previousValue = r[12];
    }
    HipRayBatchBackend be(desc, 4096);
    if (!be.ok()) { std::fprintf(stderr, "backend: %s\n", be.error().c_str()); return 3; }
    if (argc >= 8) {
        float lo[3] = {(float)std::atof(argv[5]), (float)std::atof(argv[5]), (float)std::atof(argv[5])};
        float hi[3] = {(float)std::atof(argv[6]), (float)std::atof(argv[6]), (float)std::atof(argv[6])};
        if (!be.setDomain(lo, hi, (std::uint32_t)std::atol(argv[7]))) { std::fprintf(stderr, "setDomain failed\n"); return 3; }
    }
    // split into 5 uneven batches, all submitted before any is awaited
    std::atomic<int> completed{0};
    std::vector<std::future<std::vector<RaySegmentReply>>> futs;
    std::vector<std::pair<long, long>> ranges;
    const long cuts[6] = {0, n / 7, n / 3, n / 2, (4 * n) / 5, n};
    for (int b = 0; b < 5; ++b) {
        std::vector<RaySegmentRequest> part(reqs.begin() + cuts[b], reqs.begin() + cuts[b + 1]);
        futs.push_back(be.submitBatch(part, [&completed] { completed.fetch_add(1); }));
        ranges.push_back({cuts[b], cuts[b + 1]});
    }
    // an invalid request (jitter outside [0,1)) must come back valid = false
    RaySegmentRequest bad = reqs.empty() ? RaySegmentRequest{} : reqs[0];
    bad.jitter = 2.0f;
    auto badFut = be.submitBatch({bad});
    std::vector<float> out((size_t)n * 8);
    for (int b = 4; b >= 0; --b) {  // reverse completion order
        auto rep = futs[(size_t)b].get();
        if ((long)rep.size() != ranges[(size_t)b].second - ranges[(size_t)b].first) { std::fprintf(stderr, "size\n"); return 4; }
        for (long j = 0; j < (long)rep.size(); ++j) {
            float *o = &out[(size_t)(ranges[(size_t)b].first + j) * 8];
            const RaySegmentResult &r = rep[(size_t)j].result;
            o[0] = r.t; o[1] = r.value; o[2] = r.gradient[0]; o[3] = r.gradient[1]; o[4] = r.gradient[2];
            std::memcpy(&o[5], &r.outcome, 4);
            std::memcpy(&o[6], &r.evaluations, 4);
            o[7] = rep[(size_t)j].valid ? 1.0f : 0.0f;
        }
    }
    auto badRep = badFut.get();
    if (badRep.size() != 1 || badRep[0].valid) { std::fprintf(stderr, "invalid request was served\n"); return 5; }
    // onComplete runs after publication, so wait briefly for the last callback
    for (int spin = 0; spin < 1000 && completed.load() < 5; ++spin) std::this_thread::sleep_for(std::chrono::milliseconds(1));
    if (completed.load() != 5) { std::fprintf(stderr, "onComplete count %d\n", completed.load()); return 6; }
    f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(out.data(), 4, out.size(), f) != out.size()) return 7;
    std::fclose(f);
    BackendStats s = be.stats();
    std::printf("backend_selftest: %ld requests, %llu batches, %llu rejected\n", n, (unsigned long long)s.batches,
                (unsigned long long)s.rejected);
    return 0;
}
