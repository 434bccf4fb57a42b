// memcpy of a list of host segments split over a small pool of CPU threads (the caller takes part). Used by the
// host-pointer entry points to fill and drain their pinned staging chunks. Plain C++ with no HIP dependency, so the
// CPU test suite builds it on its own under -fsanitize=thread.
//
// One list runs at a time (call_mu_). A list's parts are built in a local vector and installed in the same critical
// section of mu_ that resets next_ and done_ and advances gen_, so a pool thread that is still leaving the previous
// list sees either the old, finished list (nothing left to take) or the whole new one, never a half-built one. copy()
// returns only when every part of its list has been copied: no thread is inside a memcpy of an old part when the next
// list is installed.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace kbwt {

struct CopySeg {
    uint8_t *dst;
    const uint8_t *src;
    size_t n;
};

class CopyPool {
public:
    static constexpr size_t PART = (size_t)512 << 10;

    explicit CopyPool(unsigned threads = 0) : want_(threads) {}

    void copy(const std::vector<CopySeg> &segs) {
        std::vector<CopySeg> parts;
        for (const CopySeg &g : segs)
            for (size_t o = 0; o < g.n; o += PART) parts.push_back({g.dst + o, g.src + o, std::min(PART, g.n - o)});
        if (parts.size() < 2) {
            for (const CopySeg &g : parts) memcpy(g.dst, g.src, g.n);
            return;
        }
        std::lock_guard<std::mutex> one(call_mu_);
        start();
        size_t total;
        {
            std::lock_guard<std::mutex> lk(mu_);
            parts_.swap(parts);
            total = parts_.size();
            next_ = 0;
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return done_ == total; });
    }

    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }

private:
    void start() {   // under call_mu_
        if (!th_.empty()) return;
        unsigned k = want_;
        if (k == 0) {
            const unsigned hw = std::thread::hardware_concurrency();
            k = hw > 2 ? std::min(7u, hw - 1) : 1u;
        }
        for (unsigned i = 0; i < k; i++) th_.emplace_back([this] { loop(); });
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
        }
    }
    void work() {   // take parts of the installed list until none is left
        for (;;) {
            CopySeg g;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (next_ >= parts_.size()) return;
                g = parts_[next_++];
            }
            memcpy(g.dst, g.src, g.n);
            std::lock_guard<std::mutex> lk(mu_);
            if (++done_ == parts_.size()) done_cv_.notify_all();
        }
    }
    unsigned want_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    std::vector<std::thread> th_;
    std::vector<CopySeg> parts_;
    size_t next_ = 0, done_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace kbwt
