// Gathers concurrent single-block calls into batches. kolmo's block workers call kolmo_sais_bwt_aux and
// kolmo_sais_unbwt_aux from several threads at once (src/core/compress.c:231-256, src/core/decode.c:404-410). A call
// queues its block; a caller that finds a free lane and no batch still staging its inputs takes every queued block up
// to the batch limits, runs the forward blocks as one batch and the inverse blocks as another on that lane, and wakes
// the others. Each block's result is the single-block result, so batching changes nothing but the launch count.
//
// Up to `lanes` batches are in flight at once (two by default), so that one batch's host work overlaps another's
// device work. The runner tells the combiner through Intake::done() when its batch's inputs are staged; only then may
// the next batch form, so the blocks that queue meanwhile go into it together instead of into batches of one.
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
constexpr int COMBINE_LANES = 2;

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

class Combiner;

// Handed to the runner with its batch: done() says that the batch's inputs are staged, so the next batch may form.
// Calls after the first do nothing; staged() tells whether it has come (the blocks a failed batch runs again one by
// one stage after it).
class Intake {
public:
    explicit Intake(Combiner *c) : c_(c) {}
    inline void done();
    bool staged() const { return over_; }

private:
    Combiner *c_;
    bool over_ = false;
};

class Combiner {
public:
    explicit Combiner(int lanes = COMBINE_LANES) : lanes_(lanes < 1 ? 1 : lanes > 8 ? 8 : lanes) {
        free_ = (1u << lanes_) - 1u;
    }

    // Runner: void run(std::vector<Pending *> &batch, bool fwd, int lane, Intake &in) runs one batch of one kind on
    // lane `lane` (0 <= lane < lanes; no other batch holds it meanwhile), calls in.done() once the batch's inputs are
    // staged (optional: the combiner calls it after the runner returns), and sets every block's rc and err; it must
    // not throw. Returns p's rc; err gets p's message (empty when rc == 0).
    template <class Runner> int32_t submit(Pending &p, Runner &run, std::string &err) {
        std::unique_lock<std::mutex> lk(mu_);
        try {
            q_.push_back(&p);
        } catch (...) {
            err = "host allocation failed";
            return -2;
        }
        while (!p.done) {
            if (intake_ || free_ == 0 || q_.empty()) {
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
                cv_.notify_all();
                continue;
            }
            q_.erase(q_.begin(), q_.begin() + (long)k);
            batches_++;
            blocks_ += taken.size();
            int lane = 0;
            while (!(free_ & (1u << lane))) lane++;
            free_ &= ~(1u << lane);
            intake_ = true;
            lk.unlock();
            Intake in(this);
            const size_t re = run_isolated(fwd, true, run, lane, in) + run_isolated(inv, false, run, lane, in);
            in.done();
            lk.lock();
            retried_ += re;
            for (Pending *q : taken) q->done = true;
            free_ |= 1u << lane;
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
    friend class Intake;
    void intake_over() {
        std::lock_guard<std::mutex> lk(mu_);
        intake_ = false;
        cv_.notify_all();
    }

    // Returns the blocks run again on their own.
    template <class Runner>
    static size_t run_isolated(std::vector<Pending *> &b, bool fwd, Runner &run, int lane, Intake &in) {
        if (b.empty()) return 0;
        run(b, fwd, lane, in);
        if (b.size() < 2) return 0;
        bool failed = false;
        for (Pending *q : b) failed |= q->rc != 0;
        if (!failed) return 0;
        for (Pending *q : b) {   // every block again on its own: a failure stays with the block it belongs to
            q->rc = 0;
            q->err.clear();
            try {
                std::vector<Pending *> one(1, q);
                run(one, fwd, lane, in);
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
    int lanes_;
    uint32_t free_;            // lanes no batch holds
    bool intake_ = false;      // a batch is forming or staging its inputs
    uint64_t batches_ = 0, blocks_ = 0, retried_ = 0;
};

inline void Intake::done() {
    if (over_) return;
    over_ = true;
    c_->intake_over();
}

}  // namespace kbwt
