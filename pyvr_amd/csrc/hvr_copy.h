// hvr_copy.h -- host copies of whole frames, split over worker threads.
// Plain C++17 (no HIP), so that a native test driver can run it under
// ThreadSanitizer (tests/native/copy_check.cpp).
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>

namespace hvr {

// Host copies of whole frames from pinned memory (4 MiB at 1024^2: on one
// core the copy is bound by that core's memory bandwidth), split between the
// calling thread and three worker threads.  The workers start on first use
// (detached; they wait on a condition variable between copies) and serve one
// copy at a time.  In a process forked after they started (whose copies of
// the locks may be held by threads that do not exist there), and if the
// workers cannot be started, copies stay on the calling thread.
class CopyPool {
public:
    static CopyPool &get()
    {
        static CopyPool *pool = new CopyPool();   // never destroyed: the workers outlive static destructors
        return *pool;
    }
    void copy(uint8_t *dst, const uint8_t *src, size_t n)
    {
        const pid_t owner = owner_.load(std::memory_order_acquire);
        if (n < ((size_t)1 << 20) || (owner != 0 && owner != getpid())) {
            memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> one(call_mu_);
        const size_t part = ((n + kWorkers) / (kWorkers + 1) + 63) & ~(size_t)63;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!start()) {
                memcpy(dst, src, n);
                return;
            }
            for (int w = 0; w < kWorkers; ++w) {
                const size_t b = std::min(n, (size_t)(w + 1) * part), e = std::min(n, (size_t)(w + 2) * part);
                job_[w] = Job{dst + b, src + b, e - b};
            }
            left_ = kWorkers;
            ++gen_;
        }
        cv_.notify_all();
        memcpy(dst, src, std::min(n, part));
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return left_ == 0; });
    }

private:
    static constexpr int kWorkers = 3;
    struct Job {
        uint8_t *dst;
        const uint8_t *src;
        size_t n;
    };
    // with mu_ held: start the workers once; false when they could not all
    // be started (those that did wait for a generation that never comes)
    bool start()
    {
        if (owner_.load(std::memory_order_relaxed) == 0) {
            int started = 0;
            try {
                for (; started < kWorkers; ++started) std::thread([this, w = started] { run(w); }).detach();
            } catch (...) {
            }
            ok_ = started == kWorkers;
            owner_.store(getpid(), std::memory_order_release);
        }
        return ok_;
    }
    void run(int w)
    {
        uint64_t seen = 0;
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                j = job_[w];
            }
            if (j.n) memcpy(j.dst, j.src, j.n);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_.notify_one();
        }
    }
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    Job job_[kWorkers] = {};
    int left_ = 0;
    uint64_t gen_ = 0;
    bool ok_ = false;
    std::atomic<pid_t> owner_{0};   // the process that started the workers
};

}  // namespace hvr
