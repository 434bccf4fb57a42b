// Gathers concurrent single-block calls into batches. kolmo's block workers call kolmo_sais_bwt_aux and
// kolmo_sais_unbwt_aux from several threads at once (src/core/compress.c:231-256, src/core/decode.c:404-410). A call
// queues its block; a caller that finds no batch running takes every queued block up to the batch limits, runs the
// forward blocks as one batch and the inverse blocks as another, and wakes the others. Each block's result is the
// single-block result, so batching changes nothing but the launch count.
//
// Failures stay per block: when a batch of several blocks fails, each of its blocks runs again on its own, so one
// block's failure is not every caller's, and a batch that fails as a whole does not fail blocks that succeed alone.
// A queued block longer than one batch may hold is refused with KBWT_ERR_NOMEM instead of staying queued.
//
// Plain C++ with no HIP dependency: the CPU test suite drives it with a fake batch runner under -fsanitize=thread.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace kbwt {

constexpr int32_t BATCH_MAX_BLOCKS = 65535;
constexpr int64_t BATCH_MAX_TOTAL = (int64_t)INT32_MAX - 64;   // packed length of one batch (so of one block too)

struct Pending {
    bool fwd = true;
    const uint8_t *T = nullptr;
    uint8_t *U = nullptr;
    int32_t n = 0, r = 0;          // r: the sampling interval (forward) or r of the inverse
    int32_t *I = nullptr;          // forward: the sampled indexes out
    const int32_t *Iin = nullptr;  // inverse: the sampled indexes in
    int32_t threads = 1;           // the caller's thread count (used by a CPU path only)
    int32_t rc = 0;
    bool done = false;
    std::string err;
};

class Combiner {
public:
    // Runner: void run(std::vector<Pending *> &batch, bool fwd) runs one batch of one kind and sets every block's rc
    // and err; it must not throw. Returns p's rc; err gets p's message (empty when rc == 0).
    template <class Runner> int32_t submit(Pending &p, Runner &run, std::string &err) {
        std::unique_lock<std::mutex> lk(mu_);
        try {
            q_.push_back(&p);
        } catch (...) {
            err = "host allocation failed";
            return -2;
        }
        while (!p.done) {
            if (busy_) {
                cv_.wait(lk);
                continue;
            }
            std::vector<Pending *> fwd, inv, taken;
            try {
                fwd.reserve(q_.size());
                inv.reserve(q_.size());
                taken.reserve(q_.size());
            } catch (...) {
                q_.erase(std::remove(q_.begin(), q_.end(), &p), q_.end());
                err = "host allocation failed";
                return -2;
            }
            busy_ = true;
            int64_t tf = 0, ti = 0;
            size_t k = 0;
            for (; k < q_.size(); k++) {
                Pending *q = q_[k];
                int64_t &t = q->fwd ? tf : ti;
                std::vector<Pending *> &v = q->fwd ? fwd : inv;
                if ((int64_t)v.size() == BATCH_MAX_BLOCKS || t + q->n > BATCH_MAX_TOTAL) break;
                t += q->n;
                v.push_back(q);
                taken.push_back(q);
            }
            if (k == 0) {   // the head alone is longer than a batch: refuse it rather than leave it queued for ever
                Pending *q = q_.front();
                q_.erase(q_.begin());
                q->rc = -2;
                q->err = "block longer than one batch may hold";
                q->done = true;
                busy_ = false;
                cv_.notify_all();
                continue;
            }
            q_.erase(q_.begin(), q_.begin() + (long)k);
            batches_++;
            blocks_ += taken.size();
            lk.unlock();
            const size_t re = run_isolated(fwd, true, run) + run_isolated(inv, false, run);
            lk.lock();
            retried_ += re;
            for (Pending *q : taken) q->done = true;
            busy_ = false;
            cv_.notify_all();
        }
        try {
            err = p.rc ? p.err : std::string();
        } catch (...) {
        }
        return p.rc;
    }

    // Batches run, blocks taken, and blocks run again on their own after their batch failed, so far.
    void counts(uint64_t &batches, uint64_t &blocks, uint64_t &retried) {
        std::lock_guard<std::mutex> lk(mu_);
        batches = batches_;
        blocks = blocks_;
        retried = retried_;
    }

private:
    // Returns the blocks run again on their own.
    template <class Runner> static size_t run_isolated(std::vector<Pending *> &b, bool fwd, Runner &run) {
        if (b.empty()) return 0;
        run(b, fwd);
        if (b.size() < 2) return 0;
        bool failed = false;
        for (Pending *q : b) failed |= q->rc != 0;
        if (!failed) return 0;
        for (Pending *q : b) {   // every block again on its own: a failure stays with the block it belongs to
            q->rc = 0;
            q->err.clear();
            try {
                std::vector<Pending *> one(1, q);
                run(one, fwd);
            } catch (...) {
                q->rc = -2;
                q->err = "host allocation failed";
            }
        }
        return b.size();
    }

    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<Pending *> q_;
    bool busy_ = false;
    uint64_t batches_ = 0, blocks_ = 0, retried_ = 0;
};

}  // namespace kbwt
