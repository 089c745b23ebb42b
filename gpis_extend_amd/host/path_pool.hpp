// path_pool.hpp — a wavefront path pool over HipRayBatchBackend: the host side of the
// integrator's GPU segment contract, in the shape of PathTracer's path pool
// (PathTraceIntegrator.cpp:379-650) and its resumption protocol (advanceSample /
// resumeGpis, PathTracer.cpp:294-438).
//
// Many paths are in flight at once.  Each one advances on the host until it needs a
// GPIS segment (prepareGpisSegment: seeds from pixelSampleSegment, the jitter draw,
// the conditioning state), parks its request, and is resumed with the reply
// (finishGpisSegment / resumeDistance: t, hit, value and the endpoint gradient, which
// the reply carries so the host needs no CPU gradient).  A round gathers the parked
// requests of every live path, submits them with submitBatch, and resumes each path
// from its reply in request order; finished paths are replaced from the queue.
//
// The scene is the one gpis_trace_paths_1d restates on the device: a 1D_sampling medium
// in the box [-box, box]^3, a mirror or Lambert BRDF phase of albedo 1 at each hit, and
// for the Lambert phase an optional point light whose light sample is a Shadow segment
// conditioned on (0, gradient).  The path sampler's draws are the same PCG streams, so
// the two integration levels produce the same paths.  (The mirror's light sample needs
// neePDF of the realization, which the segment reply does not carry: the pool refuses a
// light with the mirror phase.)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gpis_backend.hpp"

namespace gpis_host {

struct PathRequest {
    float origin[3] = {0, 0, 0}, direction[3] = {0, 0, 1};
    float nearT = 0.0f, farT = 0.0f, jitter = 0.0f;
    std::uint32_t pixelSampleSegment[4] = {0, 0, 0, 0};  // (px, py, spp, bounce) of the first flight
    bool firstScatter = true;
    float previousValue = 0.0f, previousGradient[3] = {0, 0, 0};
};

struct PathResult {
    float t0 = 0.0f, gradient0[3] = {0, 0, 0};
    std::uint32_t outcome0 = GPIS_INVALID;
    std::uint32_t flights = 0, hits = 0, shadows = 0, visible = 0;
    float nee = 0.0f;  // sum of f / |light - p|^2 over unoccluded light samples
    float tLast = 0.0f;
    std::uint32_t outcomeLast = GPIS_INVALID;
};

struct PathPoolSettings {
    int maxBounces = 4;              // up to 1 + maxBounces flights per path
    int phase = GPIS_PHASE_MIRROR;   // GPIS_PHASE_MIRROR or GPIS_PHASE_LAMBERT
    std::uint32_t bounceSeed = 7;
    float box = 1.0f;
    bool useLight = false;
    float light[3] = {0, 0, 0};
    std::uint32_t sceneSeed = 0;
    std::size_t poolSize = 65536;    // paths in flight
    std::size_t batchSize = 16384;   // requests per submitBatch
};

struct PathPoolStats {
    std::uint64_t rounds = 0, batches = 0, segments = 0;
};

class PathPool {
public:
    // `desc` is the medium's descriptor (seeds and the bounce limit come from it);
    // `backend` traces the segments.
    PathPool(HipRayBatchBackend &backend, const gpis_field_desc &desc, const PathPoolSettings &settings);
    // Trace every path; false (with error()) when a setting is refused or a reply is invalid.
    bool run(const std::vector<PathRequest> &paths, std::vector<PathResult> &results);
    const std::string &error() const { return error_; }
    const PathPoolStats &stats() const { return stats_; }

private:
    struct State;
    void start(State &s, std::uint32_t index, const PathRequest &p);
    void resume(State &s, const RaySegmentReply &reply);
    void bounce(State &s);
    void beginSegment(State &s, bool shadow, float jitter);

    HipRayBatchBackend &backend_;
    gpis_field_desc desc_;
    PathPoolSettings set_;
    std::vector<PathResult> *out_ = nullptr;
    PathPoolStats stats_;
    std::string error_;
};

}  // namespace gpis_host
