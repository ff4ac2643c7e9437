// sarf_hostcopy.hpp -- parallel memcpy on persistent host threads.
//
// Used by the host-buffer focus path to move the caller's pageable arrays to
// and from pinned staging: one thread copying pageable memory moves a few
// GB/s, so the copy is split into parts taken by the calling thread and a
// few persistent workers.  Plain C++ (no HIP), so it also builds and runs
// under a host sanitizer on its own.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace sarf {

// A job is a generation number; a part is claimed by a compare-and-swap on
// (generation << 32 | part), so a worker still looking at an older job can
// never take a part of a newer one.
class HostCopier {
  public:
    // nthreads copying in all (the caller included); parts of `part` bytes.
    explicit HostCopier(int nthreads, size_t part = size_t(1) << 20) : part_(part) {
        for (int i = 0; i + 1 < nthreads; i++) th_.emplace_back([this] { work(); });
    }
    ~HostCopier() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }
    // Blocking parallel memcpy of n bytes.
    void copy(void *dst, const void *src, size_t n) {
        if (th_.empty() || n <= 2 * part_) {
            std::memcpy(dst, src, n);
            return;
        }
        Job j{(char *)dst, (const char *)src, n, (int)std::min<size_t>(256, (n + part_ - 1) / part_)};
        uint32_t g;
        {
            std::lock_guard<std::mutex> lk(m_);
            g = ++gen_;
            job_ = j;
            left_.store(j.parts);
            ticket_.store((uint64_t)g << 32);
        }
        cv_.notify_all();
        chew(j, g);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return left_.load() == 0; });
    }

  private:
    struct Job {
        char *dst;
        const char *src;
        size_t n;
        int parts;
    };
    void chew(const Job &j, uint32_t g) {
        for (;;) {
            uint64_t t = ticket_.load();
            int p;
            do {
                if ((uint32_t)(t >> 32) != g) return;  // a newer job
                p = (int)(uint32_t)t;
                if (p >= j.parts) return;
            } while (!ticket_.compare_exchange_weak(t, t + 1));
            const size_t b = j.n * p / j.parts, e = j.n * (p + 1) / j.parts;
            std::memcpy(j.dst + b, j.src + b, e - b);
            if (left_.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(m_);
                done_.notify_all();
            }
        }
    }
    void work() {
        uint32_t seen = 0;
        for (;;) {
            Job j;
            uint32_t g;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = g = gen_;
                j = job_;
            }
            chew(j, g);
        }
    }
    const size_t part_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    Job job_{};
    uint32_t gen_ = 0;
    bool stop_ = false;
    std::atomic<uint64_t> ticket_{0};
    std::atomic<int> left_{0};
};

}  // namespace sarf
