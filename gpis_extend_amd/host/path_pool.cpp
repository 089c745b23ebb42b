// path_pool.cpp — PathPool: wavefront paths over HipRayBatchBackend (see path_pool.hpp).
// Float arithmetic is written in the reference's order and compiled without contraction,
// so the host's hit-time phase rounds as the device's does.
#include "path_pool.hpp"

#include <cmath>
#include <deque>
#include <future>

#include "../csrc/gpis_device.hpp"

#pragma clang fp contract(off)

namespace gpis_host {

namespace {

struct V3 {
    float x, y, z;
};
// Vec::dot, Vec::normalized (Vec.hpp:135-197): sequential sums, multiply by 1/length.
float dot(const V3 &a, const V3 &b) {
    float s = a.x * b.x;
    s += a.y * b.y;
    s += a.z * b.z;
    return s;
}
V3 normalized(const V3 &v) {
    const float inv = 1.0f / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
// TangentFrame(n) (TangentFrame.hpp:74-85), toLocal / toGlobal.
struct Frame {
    V3 t, b, n;
};
Frame tangentFrame(V3 n) {
    Frame F;
    F.n = normalized(n);
    const float sign = std::copysign(1.0f, F.n.z);
    const float a = -1.0f / (sign + F.n.z);
    const float b = F.n.x * F.n.y * a;
    F.t = normalized({1.0f + sign * F.n.x * F.n.x * a, sign * b, -sign * F.n.x});
    F.b = normalized({b, sign + F.n.y * F.n.y * a, -F.n.y});
    return F;
}
V3 toLocal(const Frame &F, const V3 &p) { return {dot(F.t, p), dot(F.b, p), dot(F.n, p)}; }
V3 toGlobal(const Frame &F, const V3 &p) {
    const V3 a = {F.t.x * p.x, F.t.y * p.x, F.t.z * p.x};
    const V3 b = {F.b.x * p.y, F.b.y * p.y, F.b.z * p.y};
    const V3 c = {F.n.x * p.z, F.n.y * p.z, F.n.z * p.z};
    return {(a.x + b.x) + c.x, (a.y + b.y) + c.y, (a.z + b.z) + c.z};
}
// BRDFPhaseFunction::setEventIsectInfo: frame of Vec3f(aniso.normalized()), the gradient
// normalised in double.
Frame hitFrame(const float g[3]) {
    const double x = g[0], y = g[1], z = g[2];
    double l2 = x * x;
    l2 += y * y;
    l2 += z * z;
    const double inv = 1.0 / std::sqrt(l2);
    return tangentFrame({(float)(x * inv), (float)(y * inv), (float)(z * inv)});
}
float boxExit(const V3 &o, const V3 &d, float b) {
    const float tx = d.x != 0.f ? std::fmax((b - o.x) / d.x, (-b - o.x) / d.x) : 3.4e38f;
    const float ty = d.y != 0.f ? std::fmax((b - o.y) / d.y, (-b - o.y) / d.y) : 3.4e38f;
    const float tz = d.z != 0.f ? std::fmax((b - o.z) / d.z, (-b - o.z) / d.z) : 3.4e38f;
    return std::fmax(std::fmin(tx, std::fmin(ty, tz)), 0.f);
}
// Angle.hpp:8-14 in float: PI = 3.1415926536f, TWO_PI = PI * 2.0f, INV_PI = 1.0f / PI
constexpr float kPi = 3.1415926536f;

}  // namespace

// One path: PathState's fields that the medium branch uses.
struct PathPool::State {
    std::uint32_t index = 0;
    bool live = false;
    std::uint32_t px = 0, py = 0, spp = 0, w0 = 0;
    int b = 0;                        // flight index within the path
    bool shadow = false;              // the parked segment is a light sample's shadow segment
    RaySegmentRequest req;            // the parked segment (prepareGpisSegment)
    bool parked = false;
    V3 hit = {0, 0, 0}, grad = {0, 0, 0}, dir = {0, 0, 0};
    float neeF = 0.0f, neeRsq = 1.0f;
};

PathPool::PathPool(HipRayBatchBackend &backend, const gpis_field_desc &desc, const PathPoolSettings &settings)
    : backend_(backend), desc_(desc), set_(settings) {}

// prepareGpisSegment's seeds (SparseConvolutionNoiseMedium.cpp:235-238, computeSeed
// SparseConvolutionNoise.cpp:57-66) for pixelSampleSegment (px, py, spp, w).
static void segmentSeeds(const gpis_field_desc &d, std::uint32_t sceneSeed, std::uint32_t px, std::uint32_t py,
                         std::uint32_t spp, std::uint32_t w, std::uint32_t &noise, std::uint32_t &normal) {
    const std::uint32_t h = gpis::xxh4(px, py, spp, w), hs = gpis::xxh1(sceneSeed);
    normal = h + hs + 1u;
    noise = d.global_seed;
    if (!d.single_realization) noise += (d.ctxt == GPIS_CTX_GLOBAL ? gpis::xxh4(px, py, spp, 0u) : h) + hs;
}

void PathPool::beginSegment(State &s, bool shadow, float jitter) {
    s.shadow = shadow;
    s.req.purpose = shadow ? RaySegmentRequest::Purpose::Shadow : RaySegmentRequest::Purpose::Path;
    s.req.jitter = jitter;
    segmentSeeds(desc_, set_.sceneSeed, s.px, s.py, s.spp, s.w0 + (std::uint32_t)s.b, s.req.noiseSeed,
                 s.req.normalSeed);
    s.parked = true;
}

void PathPool::start(State &s, std::uint32_t index, const PathRequest &p) {
    s = State();
    s.index = index;
    s.live = true;
    s.px = p.pixelSampleSegment[0]; s.py = p.pixelSampleSegment[1];
    s.spp = p.pixelSampleSegment[2]; s.w0 = p.pixelSampleSegment[3];
    PathResult &r = (*out_)[index];
    r = PathResult();
    for (int a = 0; a < 3; ++a) {
        s.req.origin[a] = p.origin[a];
        s.req.direction[a] = p.direction[a];
        s.req.previousGradient[a] = p.previousGradient[a];
    }
    s.req.nearT = p.nearT;
    s.req.farT = std::isfinite(p.farT) ? p.farT : p.nearT + 2000.0f;
    s.req.firstScatter = p.firstScatter;
    s.req.previousValue = p.previousValue;
    // beginDistance (GaussianProcessMedium.cpp:244-265): the medium's bounce limit, an empty interval
    if ((int)s.w0 >= desc_.max_bounce || s.req.farT <= s.req.nearT) {
        r.flights = 1;
        r.t0 = r.tLast = s.req.farT;
        r.outcome0 = r.outcomeLast = (int)s.w0 >= desc_.max_bounce ? (GPIS_FAIL | (GPIS_FAIL_BOUNCE_LIMIT << 8))
                                                                   : GPIS_EXIT;
        s.live = false;
        return;
    }
    beginSegment(s, false, p.jitter);
}

void PathPool::bounce(State &s) {
    PathResult &r = (*out_)[s.index];
    if (s.b >= set_.maxBounces) { s.live = false; return; }
    const float g[3] = {s.grad.x, s.grad.y, s.grad.z};
    const Frame F = hitFrame(g);
    const V3 wi = normalized(toLocal(F, {-s.dir.x, -s.dir.y, -s.dir.z}));
    gpis::Pcg rng;
    rng.set((std::uint64_t)(std::uint32_t)(gpis::xxh4(s.index, (std::uint32_t)(s.b + 1), set_.bounceSeed, 0u) + 1u));
    const float u1 = gpis::u01(rng.next()), u2 = gpis::u01(rng.next()), jit = gpis::u01(rng.next());
    V3 wl;
    if (set_.phase == GPIS_PHASE_MIRROR) {
        wl = {-wi.x, -wi.y, wi.z};  // MirrorBsdf::sample
    } else {
        if (!(wi.z > 0.0f)) { s.live = false; return; }  // LambertBsdf::sample refuses
        const float phi = u1 * (kPi * 2.0f), rr = std::sqrt(u2);  // SampleWarp::cosineHemisphere
        wl = {std::cos(phi) * rr, std::sin(phi) * rr, std::sqrt(std::fmax(1.0f - u2, 0.0f))};
    }
    const V3 w = normalized(toGlobal(F, wl));
    ++s.b;
    for (int a = 0; a < 3; ++a) {
        s.req.origin[a] = (&s.hit.x)[a];
        s.req.direction[a] = (&w.x)[a];
        s.req.previousGradient[a] = (&s.grad.x)[a];
    }
    s.req.nearT = 0.0f;
    s.req.farT = boxExit(s.hit, w, set_.box);
    s.req.firstScatter = false;
    s.req.previousValue = 0.0f;
    if ((int)(s.w0 + (std::uint32_t)s.b) >= desc_.max_bounce || s.req.farT <= s.req.nearT) {
        ++r.flights;
        r.tLast = s.req.farT;
        r.outcomeLast = s.req.farT <= s.req.nearT && (int)(s.w0 + (std::uint32_t)s.b) < desc_.max_bounce
                            ? GPIS_EXIT : (GPIS_FAIL | (GPIS_FAIL_BOUNCE_LIMIT << 8));
        s.live = false;
        return;
    }
    beginSegment(s, false, jit);
}

// resumeGpis + the rest of advanceSample's medium branch for one reply.
void PathPool::resume(State &s, const RaySegmentReply &reply) {
    PathResult &r = (*out_)[s.index];
    s.parked = false;
    const std::uint32_t oc = reply.result.outcome & 0xFFu;
    if (s.shadow) {
        // resumeTransmittance (GaussianProcessMedium.cpp:394-421): any crossing occludes
        ++r.shadows;
        if (oc == GPIS_EXIT) {
            ++r.visible;
            r.nee += s.neeF / s.neeRsq;
        }
        bounce(s);
        return;
    }
    ++r.flights;
    if (s.b == 0) {
        r.t0 = reply.result.t;
        for (int a = 0; a < 3; ++a) r.gradient0[a] = reply.result.gradient[a];
        r.outcome0 = reply.result.outcome;
    }
    r.tLast = reply.result.t;
    r.outcomeLast = reply.result.outcome;
    if (oc != GPIS_HIT) { s.live = false; return; }
    ++r.hits;
    // finishDistance: sample.p = ray.pos() + sample.t * ray.dir() in float
    const float t = reply.result.t;
    s.dir = {s.req.direction[0], s.req.direction[1], s.req.direction[2]};
    s.hit = {s.req.origin[0] + t * s.dir.x, s.req.origin[1] + t * s.dir.y, s.req.origin[2] + t * s.dir.z};
    s.grad = {reply.result.gradient[0], reply.result.gradient[1], reply.result.gradient[2]};
    if (set_.useLight) {
        // volumeLightSample (TraceBase.cpp:798-839) with Point::sampleDirect (Point.cpp:98-106)
        V3 dl = {set_.light[0] - s.hit.x, set_.light[1] - s.hit.y, set_.light[2] - s.hit.z};
        s.neeRsq = dot(dl, dl);
        const float dist = std::sqrt(s.neeRsq);
        dl = {dl.x / dist, dl.y / dist, dl.z / dist};
        const float g[3] = {s.grad.x, s.grad.y, s.grad.z};
        const Frame F = hitFrame(g);
        const V3 wi = normalized(toLocal(F, {-s.dir.x, -s.dir.y, -s.dir.z}));
        const V3 wo = normalized(toLocal(F, dl));
        s.neeF = (wi.z > 0.0f && wo.z > 0.0f) ? (1.0f / kPi) * wo.z : 0.0f;  // LambertBsdf::eval
        if (s.neeF != 0.0f) {
            for (int a = 0; a < 3; ++a) {
                s.req.origin[a] = (&s.hit.x)[a];
                s.req.direction[a] = (&dl.x)[a];
                s.req.previousGradient[a] = (&s.grad.x)[a];
            }
            s.req.nearT = 0.0f;
            s.req.farT = std::fmin(dist, boxExit(s.hit, dl, set_.box));
            s.req.firstScatter = false;
            s.req.previousValue = 0.0f;
            gpis::Pcg rng;
            rng.set((std::uint64_t)(std::uint32_t)(gpis::xxh4(s.index, (std::uint32_t)s.b, set_.bounceSeed, 1u) + 1u));
            const float jit = gpis::u01(rng.next());
            if (s.req.farT <= s.req.nearT) {
                ++r.shadows; ++r.visible;  // transmittance over an empty segment
                r.nee += s.neeF / s.neeRsq;
            } else {
                beginSegment(s, true, jit);
                return;
            }
        }
    }
    bounce(s);
}

bool PathPool::run(const std::vector<PathRequest> &paths, std::vector<PathResult> &results) {
    if (set_.phase != GPIS_PHASE_MIRROR && set_.phase != GPIS_PHASE_LAMBERT) { error_ = "unknown phase"; return false; }
    if (set_.useLight && set_.phase == GPIS_PHASE_MIRROR) {
        error_ = "the mirror's light sample needs neePDF, which segment replies do not carry";
        return false;
    }
    if (!backend_.ok()) { error_ = "backend disabled: " + backend_.error(); return false; }
    if (set_.batchSize == 0) { error_ = "batchSize must be positive"; return false; }
    if (set_.poolSize == 0 && !paths.empty()) { error_ = "poolSize must be positive"; return false; }
    results.assign(paths.size(), PathResult());
    out_ = &results;
    std::vector<State> pool(std::min(set_.poolSize, paths.size()));
    std::size_t next = 0;
    for (State &s : pool) {
        while (next < paths.size()) {
            start(s, (std::uint32_t)next, paths[next]);
            ++next;
            if (s.live) break;
        }
    }
    for (;;) {
        // gather the parked segments of the live paths (one round of the coordinator)
        std::vector<std::size_t> who;
        std::vector<RaySegmentRequest> reqs;
        for (std::size_t i = 0; i < pool.size(); ++i)
            if (pool[i].live && pool[i].parked) { who.push_back(i); reqs.push_back(pool[i].req); }
        if (reqs.empty()) break;
        ++stats_.rounds;
        // every batch of the round is in flight before the first reply is awaited
        std::vector<std::future<std::vector<RaySegmentReply>>> futs;
        for (std::size_t b0 = 0; b0 < reqs.size(); b0 += set_.batchSize) {
            const std::size_t b1 = std::min(reqs.size(), b0 + set_.batchSize);
            futs.push_back(backend_.submitBatch(std::vector<RaySegmentRequest>(reqs.begin() + b0, reqs.begin() + b1)));
            ++stats_.batches;
        }
        std::size_t k = 0;
        for (auto &f : futs) {
            const std::vector<RaySegmentReply> rep = f.get();
            // replies come back in request order, one per request of the batch
            const std::size_t expect = std::min(set_.batchSize, who.size() - k);
            if (rep.size() != expect) {
                error_ = "batch returned " + std::to_string(rep.size()) + " replies for " + std::to_string(expect) +
                         " requests";
                out_ = nullptr;
                return false;
            }
            for (const RaySegmentReply &r : rep) {
                State &s = pool[who[k++]];
                if (!r.valid) {
                    error_ = "invalid reply for path " + std::to_string(s.index);
                    out_ = nullptr;
                    return false;
                }
                resume(s, r);
                ++stats_.segments;
                // a finished path's slot takes the next queued path
                while (!s.live && next < paths.size()) {
                    start(s, (std::uint32_t)next, paths[next]);
                    ++next;
                }
            }
        }
    }
    out_ = nullptr;
    return true;
}

}  // namespace gpis_host
