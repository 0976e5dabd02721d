// host_stage.h -- moving a field between a caller's (pageable) host array and HBM at the
// pinned-DMA rate: the field goes in chunks through two pinned buffers, a pool of host
// threads filling (or draining) one buffer while the DMA engine copies the other.
//
// A pageable hipMemcpy stages through the runtime's own buffers with one host thread
// (28 GB/s to the device, 17 GB/s back at 273 MB); pinned DMA runs at 57 GB/s both ways, and
// several threads copy host memory faster than one, so the two overlap to near the DMA rate.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace eddy {

// memcpy split over a persistent set of threads (the caller takes the first share)
class CopyPool {
  public:
    explicit CopyPool(int nthreads) : n_(std::max(1, nthreads)) {
        try {
            for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { worker(t); });
        } catch (...) {  // join what started before passing the failure on
            shutdown();
            throw;
        }
    }
    ~CopyPool() { shutdown(); }
    void copy(void* dst, const void* src, size_t len) {
        if (n_ == 1 || len < (size_t)4 << 20) {
            std::memcpy(dst, src, len);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            len_ = len;
            left_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        share(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
    }

  private:
    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_)
            if (t.joinable()) t.join();
        th_.clear();
    }
    void share(int t) {  // 4 KiB-aligned slices
        const size_t pages = (len_ + 4095) / 4096;
        const size_t a = std::min(len_, pages * t / n_ * 4096), b = std::min(len_, pages * (t + 1) / n_ * 4096);
        if (b > a) std::memcpy(dst_ + a, src_ + a, b - a);
    }
    void worker(int t) {
        long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            share(t);
            {
                std::lock_guard<std::mutex> lk(mu_);
                --left_;
            }
            done_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t len_ = 0;
    int left_ = 0;
    long gen_ = 0;
    bool stop_ = false;
};

class HostStage {
  public:
    static constexpr size_t kChunk = (size_t)32 << 20;

    HostStage() : pool_((int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()))) {}
    ~HostStage() {
        for (int b = 0; b < 2; ++b) {
            if (pending_[b]) (void)hipEventSynchronize(ev_[b]);
            if (ev_[b]) (void)hipEventDestroy(ev_[b]);
            if (buf_[b]) (void)hipHostFree(buf_[b]);
        }
    }
    // 0, or the failing HIP error
    hipError_t init() {
        for (int b = 0; b < 2; ++b) {
            hipError_t e = hipHostMalloc((void**)&buf_[b], kChunk, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            e = hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    // host -> dev on st; returns once every byte has left `host` (the DMA may still run)
    hipError_t to_device(double* dev, const double* host, size_t n, hipStream_t st) {
        const size_t bytes = n * sizeof(double);
        const char* src = reinterpret_cast<const char*>(host);
        char* dst = reinterpret_cast<char*>(dev);
        int b = 0;
        for (size_t off = 0; off < bytes; off += kChunk, b ^= 1) {
            const size_t len = std::min(kChunk, bytes - off);
            hipError_t e = wait(b);
            if (e != hipSuccess) return e;
            pool_.copy(buf_[b], src + off, len);
            if ((e = hipMemcpyAsync(dst + off, buf_[b], len, hipMemcpyHostToDevice, st)) != hipSuccess ||
                (e = mark(b, st)) != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
    // dev -> host after the work already enqueued on st; returns when `host` holds the field
    hipError_t to_host(double* host, const double* dev, size_t n, hipStream_t st) {
        const size_t bytes = n * sizeof(double);
        const char* src = reinterpret_cast<const char*>(dev);
        char* dst = reinterpret_cast<char*>(host);
        const size_t nch = (bytes + kChunk - 1) / kChunk;
        auto issue = [&](size_t c) -> hipError_t {
            const int b = (int)(c & 1);
            hipError_t e = wait(b);
            if (e != hipSuccess) return e;
            const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
            if ((e = hipMemcpyAsync(buf_[b], src + off, len, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
            return mark(b, st);
        };
        hipError_t e = hipSuccess;
        for (size_t c = 0; c < std::min<size_t>(2, nch) && e == hipSuccess; ++c) e = issue(c);
        for (size_t c = 0; c < nch && e == hipSuccess; ++c) {
            const int b = (int)(c & 1);
            if ((e = wait(b)) != hipSuccess) break;
            const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
            pool_.copy(dst + off, buf_[b], len);
            if (c + 2 < nch) e = issue(c + 2);
        }
        return e;
    }

  private:
    hipError_t wait(int b) {
        if (!pending_[b]) return hipSuccess;
        pending_[b] = false;
        return hipEventSynchronize(ev_[b]);
    }
    hipError_t mark(int b, hipStream_t st) {
        const hipError_t e = hipEventRecord(ev_[b], st);
        if (e == hipSuccess) pending_[b] = true;
        return e;
    }
    char* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool pending_[2] = {false, false};
    CopyPool pool_;
};

}  // namespace eddy
