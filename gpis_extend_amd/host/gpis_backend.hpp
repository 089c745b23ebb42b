// gpis_backend.hpp — C++ host side of the GPIS flight engine: a batch backend with
// the shape of Tungsten's GpRayBatchBackend (src/core/media/GpRayBatchBackend.hpp:43-55)
// over the C ABI of include/gpis_hip.h.
//
// Contract kept from the reference:
//   * trace(request, result) blocks and returns false when the request cannot be served;
//   * submitBatch(requests, onComplete) copies the requests, returns a future of replies
//     in request order, unsupported or failed requests included with valid = false;
//   * onComplete runs once after publication, on the backend's worker thread, and must
//     not call into HIP (it only notifies a scheduler);
//   * a failed or invalid reply leaves the caller free to trace the identical prepared
//     segment on the CPU (SparseConvolutionNoiseMedium.cpp:265-347).
// Additions: the reply carries the endpoint gradient and the outcome, so the caller
// needs no CPU gradient evaluation.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpis_hip.h"

namespace gpis_host {

// Field-for-field mirror of GpRaySegmentRequest (GpRayBatchBackend.hpp:15-29).
struct RaySegmentRequest {
    enum class Purpose { Path, Shadow };
    Purpose purpose = Purpose::Path;
    float origin[3] = {0, 0, 0};
    float direction[3] = {0, 0, 1};
    float nearT = 0.0f;
    float farT = 0.0f;
    float jitter = 0.0f;
    std::uint32_t noiseSeed = 0;
    std::uint32_t normalSeed = 0;
    bool firstScatter = true;
    float previousValue = 0.0f;
    float previousGradient[3] = {0, 0, 0};
};

// GpRaySegmentResult (GpRayBatchBackend.hpp:31-35) plus gradient and outcome.
struct RaySegmentResult {
    float t = 0.0f;
    float value = 0.0f;
    bool hit = false;
    float gradient[3] = {0, 0, 0};
    std::uint32_t outcome = GPIS_INVALID;  // low byte GPIS_EXIT/HIT/FAIL/INVALID
    std::uint32_t evaluations = 0;
};

struct RaySegmentReply {
    bool valid = false;
    RaySegmentResult result;
};

struct BackendStats {
    std::uint64_t batches = 0, requests = 0, rejected = 0, deviceErrors = 0;
};

class HipRayBatchBackend {
public:
    // One context for one medium on the current HIP device. `capacity` is the largest
    // batch one submission moves at once (larger batches are split).
    explicit HipRayBatchBackend(const gpis_field_desc &desc, std::size_t capacity = 65536);
    ~HipRayBatchBackend();
    HipRayBatchBackend(const HipRayBatchBackend &) = delete;
    HipRayBatchBackend &operator=(const HipRayBatchBackend &) = delete;

    bool ok() const { return ctx_ != nullptr && !disabled_; }
    const std::string &error() const { return error_; }

    // Optional: precompute the packed impulse table of a single-realization medium.
    bool setDomain(const float lo[3], const float hi[3], std::uint32_t seed);
    // Required for a tabulated mean (mean_kind GPIS_MEAN_GRID) before the first trace:
    // the dense export of the mean's grid (the reference's makeHostMeanGrid output).
    bool setMeanGrid(const gpis_mean_grid &grid);
    // Arithmetic of 1D flights (gpis_set_precision: GPIS_PRECISION_FAST or _EXACT).
    bool setPrecision(std::int32_t mode);

    bool trace(const RaySegmentRequest &request, RaySegmentResult &result);
    std::future<std::vector<RaySegmentReply>> submitBatch(const std::vector<RaySegmentRequest> &requests,
                                                          std::function<void()> onComplete = {});
    BackendStats stats() const;

private:
    struct Job {
        std::vector<RaySegmentRequest> requests;
        std::function<void()> onComplete;
        std::promise<std::vector<RaySegmentReply>> promise;
    };
    void run();
    std::vector<RaySegmentReply> process(const std::vector<RaySegmentRequest> &requests);
    bool ensureBuffers();

    gpis_ctx *ctx_ = nullptr;
    std::size_t capacity_;
    bool oneD_;  // 1D_sampling medium: flights go to gpis_trace_flights_1d
    void *stream_ = nullptr;
    // pinned host staging and device buffers, SoA, `capacity_` entries each
    void *hostIn_ = nullptr, *hostOut_ = nullptr, *devIn_ = nullptr, *devOut_ = nullptr;
    bool disabled_ = false;
    std::string error_;
    mutable std::mutex mutex_;   // queue and statistics
    std::mutex deviceMutex_;     // the context, stream and staging buffers
    std::condition_variable cv_;
    std::deque<std::unique_ptr<Job>> queue_;
    bool stop_ = false;
    BackendStats stats_;
    std::thread worker_;
};

}  // namespace gpis_host
