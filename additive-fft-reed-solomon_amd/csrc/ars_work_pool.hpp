// ars_work_pool.hpp -- a small persistent worker pool for host-side work:
// the host-slice pipeline's row copies between pageable caller memory and
// pinned staging, and the per-column key equations of error decoding.
// HIP-free so that it builds (and is tested under ThreadSanitizer) with g++.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace ars {

// Persistent workers that run fn(i) for i in [0, count) together with the
// calling thread.  One job at a time (the pipe's mutex serialises callers).
class WorkPool {
  public:
    explicit WorkPool(unsigned workers) {
        for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }
    ~WorkPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void run(size_t count, const std::function<void(size_t)> &fn) {
        if (threads_.empty() || count <= 1) {
            for (size_t i = 0; i < count; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            count_ = count;
            next_.store(0);
            active_ = (unsigned)threads_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return active_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work() {
        for (size_t i; (i = next_.fetch_add(1)) < count_;) (*fn_)(i);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t count_ = 0;
    std::atomic<size_t> next_{0};
    unsigned active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace ars
