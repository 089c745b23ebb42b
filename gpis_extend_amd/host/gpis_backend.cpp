// gpis_backend.cpp — HipRayBatchBackend: packs copied requests into pinned SoA
// staging, moves them to HBM, runs gpis_trace_flights on the backend's stream and
// publishes replies in request order (GpRayBatchBackend contract,
// src/core/media/GpRayBatchBackend.hpp:43-55; dispatcher pattern of
// src/gpu/PathTracerSeBatchBackend.cpp:172-251).
#include "gpis_backend.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

namespace gpis_host {

namespace {
constexpr int kIn = 16;   // SoA input arrays, 4 bytes each
constexpr int kOut = 7;   // SoA output arrays, 4 bytes each

bool supported(const RaySegmentRequest &r) {
    if (r.purpose != RaySegmentRequest::Purpose::Path && r.purpose != RaySegmentRequest::Purpose::Shadow) return false;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(r.origin[i]) || !std::isfinite(r.direction[i])) return false;
    if (!std::isfinite(r.nearT) || !(r.jitter >= 0.0f && r.jitter < 1.0f)) return false;
    if (!r.firstScatter) {
        if (!std::isfinite(r.previousValue)) return false;
        for (int i = 0; i < 3; ++i)
            if (!std::isfinite(r.previousGradient[i])) return false;
    }
    return true;
}
}  // namespace

HipRayBatchBackend::HipRayBatchBackend(const gpis_field_desc &desc, std::size_t capacity)
    : capacity_(capacity), oneD_(desc.one_d != 0) {
    if (gpis_create(&desc, &ctx_) != GPIS_OK) {
        error_ = "gpis_create failed";
        ctx_ = nullptr;
        return;
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        error_ = "hipStreamCreate failed";
        disabled_ = true;
        return;
    }
    stream_ = s;
    if (!ensureBuffers()) return;
    worker_ = std::thread([this] { run(); });
}

HipRayBatchBackend::~HipRayBatchBackend() {
    {
        std::lock_guard<std::mutex> lk(mutex_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
    if (hostIn_) (void)hipHostFree(hostIn_);
    if (hostOut_) (void)hipHostFree(hostOut_);
    if (devIn_) (void)hipFree(devIn_);
    if (devOut_) (void)hipFree(devOut_);
    if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
    if (ctx_) gpis_destroy(ctx_);
}

bool HipRayBatchBackend::ensureBuffers() {
    const std::size_t inBytes = capacity_ * kIn * 4, outBytes = capacity_ * kOut * 4;
    if (hipHostMalloc(&hostIn_, inBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&hostOut_, outBytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&devIn_, inBytes) != hipSuccess || hipMalloc(&devOut_, outBytes) != hipSuccess) {
        error_ = "buffer allocation failed";
        disabled_ = true;
        return false;
    }
    return true;
}

bool HipRayBatchBackend::setDomain(const float lo[3], const float hi[3], std::uint32_t seed) {
    if (!ok()) return false;
    std::lock_guard<std::mutex> lk(deviceMutex_);
    return gpis_set_domain(ctx_, lo, hi, seed) == GPIS_OK;
}

bool HipRayBatchBackend::setMeanGrid(const gpis_mean_grid &grid) {
    if (!ok()) return false;
    std::lock_guard<std::mutex> lk(deviceMutex_);
    if (gpis_set_mean_grid(ctx_, &grid) != GPIS_OK) {
        error_ = gpis_last_error(ctx_);
        return false;
    }
    return true;
}

bool HipRayBatchBackend::setPrecision(std::int32_t mode) {
    if (!ok()) return false;
    std::lock_guard<std::mutex> lk(deviceMutex_);
    if (gpis_set_precision(ctx_, mode) != GPIS_OK) {
        error_ = gpis_last_error(ctx_);
        return false;
    }
    return true;
}

std::vector<RaySegmentReply> HipRayBatchBackend::process(const std::vector<RaySegmentRequest> &requests) {
    std::vector<RaySegmentReply> replies(requests.size());
    if (!ok()) return replies;  // all invalid: the caller traces them on the CPU
    std::vector<std::size_t> idx;
    idx.reserve(requests.size());
    for (std::size_t i = 0; i < requests.size(); ++i)
        if (supported(requests[i])) idx.push_back(i);
    hipStream_t st = (hipStream_t)stream_;
    for (std::size_t base = 0; base < idx.size(); base += capacity_) {
        const std::size_t n = std::min(capacity_, idx.size() - base);
        float *hf = static_cast<float *>(hostIn_);
        std::uint32_t *hu = static_cast<std::uint32_t *>(hostIn_);
        auto F = [&](int a, std::size_t j) -> float & { return hf[a * capacity_ + j]; };
        auto U = [&](int a, std::size_t j) -> std::uint32_t & { return hu[a * capacity_ + j]; };
        for (std::size_t j = 0; j < n; ++j) {
            const RaySegmentRequest &r = requests[idx[base + j]];
            F(0, j) = r.origin[0]; F(1, j) = r.origin[1]; F(2, j) = r.origin[2];
            F(3, j) = r.direction[0]; F(4, j) = r.direction[1]; F(5, j) = r.direction[2];
            F(6, j) = r.nearT; F(7, j) = r.farT; F(8, j) = r.jitter;
            U(9, j) = r.noiseSeed; U(10, j) = r.normalSeed; U(11, j) = r.firstScatter ? 1u : 0u;
            F(12, j) = r.previousValue;
            F(13, j) = r.previousGradient[0]; F(14, j) = r.previousGradient[1]; F(15, j) = r.previousGradient[2];
        }
        char *din = static_cast<char *>(devIn_);
        char *dout = static_cast<char *>(devOut_);
        bool good = true;
        for (int a = 0; a < kIn && good; ++a)
            good = hipMemcpyAsync(din + a * capacity_ * 4, static_cast<char *>(hostIn_) + a * capacity_ * 4, n * 4,
                                  hipMemcpyHostToDevice, st) == hipSuccess;
        gpis_ray_soa rays;
        auto DF = [&](int a) { return reinterpret_cast<const float *>(din + a * capacity_ * 4); };
        auto DU = [&](int a) { return reinterpret_cast<const std::uint32_t *>(din + a * capacity_ * 4); };
        rays.ox = DF(0); rays.oy = DF(1); rays.oz = DF(2); rays.dx = DF(3); rays.dy = DF(4); rays.dz = DF(5);
        rays.near_t = DF(6); rays.far_t = DF(7); rays.jitter = DF(8);
        rays.noise_seed = DU(9); rays.normal_seed = DU(10); rays.flags = DU(11);
        rays.prev_value = DF(12); rays.prev_gx = DF(13); rays.prev_gy = DF(14); rays.prev_gz = DF(15);
        rays.length_scale = nullptr;
        auto OF = [&](int a) { return reinterpret_cast<float *>(dout + a * capacity_ * 4); };
        auto OU = [&](int a) { return reinterpret_cast<std::uint32_t *>(dout + a * capacity_ * 4); };
        gpis_reply_soa rep{OF(0), OF(1), OF(2), OF(3), OF(4), OU(5), OU(6)};
        if (good)
            good = (oneD_ ? gpis_trace_flights_1d(ctx_, &rays, (int64_t)n, &rep, nullptr, nullptr, st)
                          : gpis_trace_flights(ctx_, &rays, (int64_t)n, &rep, st)) == GPIS_OK;
        for (int a = 0; a < kOut && good; ++a)
            good = hipMemcpyAsync(static_cast<char *>(hostOut_) + a * capacity_ * 4, dout + a * capacity_ * 4, n * 4,
                                  hipMemcpyDeviceToHost, st) == hipSuccess;
        if (good) good = hipStreamSynchronize(st) == hipSuccess;
        if (!good) {
            // a device error disables the backend: every later request falls back to the CPU
            // (PathTracerSeBatchBackend.cpp:108-110,218-229)
            disabled_ = true;
            error_ = "device error during a batch";
            {
                std::lock_guard<std::mutex> lk(mutex_);
                ++stats_.deviceErrors;
            }
            return std::vector<RaySegmentReply>(requests.size());
        }
        const float *of = static_cast<const float *>(hostOut_);
        const std::uint32_t *ou = static_cast<const std::uint32_t *>(hostOut_);
        for (std::size_t j = 0; j < n; ++j) {
            RaySegmentReply &rp = replies[idx[base + j]];
            RaySegmentResult &res = rp.result;
            res.t = of[0 * capacity_ + j];
            res.value = of[1 * capacity_ + j];
            res.gradient[0] = of[2 * capacity_ + j];
            res.gradient[1] = of[3 * capacity_ + j];
            res.gradient[2] = of[4 * capacity_ + j];
            res.outcome = ou[5 * capacity_ + j];
            res.evaluations = ou[6 * capacity_ + j];
            const std::uint32_t oc = res.outcome & 0xFFu;
            res.hit = oc == GPIS_HIT || oc == GPIS_FAIL;  // a crossing was found
            rp.valid = oc != GPIS_INVALID;
        }
    }
    return replies;
}

void HipRayBatchBackend::run() {
    for (;;) {
        std::unique_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mutex_);
            cv_.wait(lk, [this] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;
            job = std::move(queue_.front());
            queue_.pop_front();
        }
        std::vector<RaySegmentReply> replies;
        {
            std::lock_guard<std::mutex> dk(deviceMutex_);
            replies = process(job->requests);
        }
        {
            std::lock_guard<std::mutex> lk(mutex_);
            ++stats_.batches;
            stats_.requests += job->requests.size();
            for (const auto &r : replies) stats_.rejected += r.valid ? 0 : 1;
        }
        job->promise.set_value(std::move(replies));
        if (job->onComplete) job->onComplete();
    }
}

std::future<std::vector<RaySegmentReply>> HipRayBatchBackend::submitBatch(const std::vector<RaySegmentRequest> &requests,
                                                                          std::function<void()> onComplete) {
    auto job = std::make_unique<Job>();
    job->requests = requests;  // copied and retained until publication
    job->onComplete = std::move(onComplete);
    auto fut = job->promise.get_future();
    if (!worker_.joinable()) {
        job->promise.set_value(std::vector<RaySegmentReply>(requests.size()));
        if (job->onComplete) job->onComplete();
        return fut;
    }
    {
        std::lock_guard<std::mutex> lk(mutex_);
        queue_.push_back(std::move(job));
    }
    cv_.notify_one();
    return fut;
}

bool HipRayBatchBackend::trace(const RaySegmentRequest &request, RaySegmentResult &result) {
    auto f = submitBatch({request});
    auto r = f.get();
    result = r[0].result;
    return r[0].valid;
}

BackendStats HipRayBatchBackend::stats() const {
    std::lock_guard<std::mutex> lk(mutex_);
    return stats_;
}

}  // namespace gpis_host
