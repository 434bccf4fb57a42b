// Stress of kbwt::Combiner (kolmo_amd/csrc/kbwt_combine.h) with a fake batch runner, many caller threads at once.
// The runner's "forward" writes each byte plus one, its "inverse" each byte minus one. A batch that holds a poisoned
// block (first byte 0xEE) fails as a whole with -3, as a device error would; run alone, the poisoned block still
// fails and every other block succeeds. Checks: every caller gets its own block's bytes and code, failures stay with
// the poisoned blocks, a block longer than a batch may hold is refused without a hang, and batching happens. Lanes:
// no lane carries two batches at once, at most one batch stages its inputs at a time (before Intake::done()), and two
// batches do run at once (one staging while the other computes).
#include "kbwt_combine.h"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

int main(int argc, char **argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 100;
    kbwt::Combiner comb;
    std::atomic<int> bad{0}, poisoned_seen{0}, multi{0}, staging{0}, running{0}, overlapped{0};
    std::atomic<int> lane_busy[kbwt::COMBINE_LANES] = {};
    auto runner = [&](std::vector<kbwt::Pending *> &b, bool fwd, int lane, kbwt::Intake &in) {
        if (lane < 0 || lane >= kbwt::COMBINE_LANES || lane_busy[lane]++ != 0) bad++;
        if (running++ > 0) overlapped++;
        if (b.size() > 1) multi++;
        bool poison = false;
        for (auto *q : b) poison |= q->n > 0 && q->T[0] == 0xEE;
        const bool first = !in.staged();   // a batch's first run, not a block run again after its batch failed
        if (first && staging++ != 0) bad++;   // the "inputs": only one batch stages at a time
        std::this_thread::sleep_for(std::chrono::microseconds(100));
        if (first) staging--;
        in.done();
        std::this_thread::sleep_for(std::chrono::microseconds(200));
        for (auto *q : b) {
            if (poison) {
                q->rc = -3;
                q->err = "injected";
                continue;
            }
            for (int32_t i = 0; i < q->n; i++) q->U[i] = (uint8_t)(q->T[i] + (fwd ? 1 : -1));
            if (fwd) q->I[0] = q->n;
            q->rc = 0;
        }
        running--;
        lane_busy[lane]--;
    };
    std::vector<std::thread> th;
    for (int c = 0; c < 12; c++)
        th.emplace_back([&, c] {
            std::mt19937 rng(77 + c);
            for (int it = 0; it < iters; it++) {
                const int32_t n = 1 + (int32_t)(rng() % 5000);
                std::vector<uint8_t> T(n), U(n, 0);
                for (auto &x : T) x = (uint8_t)(rng() % 200);
                const bool poison = rng() % 10 == 0;
                if (poison) T[0] = 0xEE;
                kbwt::Pending p;
                p.fwd = rng() & 1;
                p.T = T.data();
                p.U = U.data();
                p.n = n;
                p.r = 2;
                int32_t I = -1;
                p.I = &I;
                p.Iin = &I;
                std::string err;
                const int32_t rc = comb.submit(p, runner, err);
                if (poison) {
                    poisoned_seen++;
                    if (rc != -3 || err != "injected") bad++;
                    continue;
                }
                if (rc != 0 || !err.empty()) { bad++; continue; }
                for (int32_t i = 0; i < n; i++)
                    if (U[i] != (uint8_t)(T[i] + (p.fwd ? 1 : -1))) { bad++; break; }
                if (p.fwd && I != n) bad++;
            }
        });
    // a block longer than a batch may hold: refused, never run, nobody stuck behind it
    th.emplace_back([&] {
        kbwt::Pending p;
        p.n = INT32_MAX;
        std::string err;
        if (comb.submit(p, runner, err) != -2) bad++;
    });
    for (auto &t : th) t.join();
    uint64_t batches = 0, blocks = 0, retried = 0;
    comb.counts(batches, blocks, retried);
    printf("combine: %llu batches for %llu blocks, %d multi-block runs, %d poisoned calls, %llu retried, %d runs "
           "beside another, %d bad\n",
           (unsigned long long)batches, (unsigned long long)blocks, multi.load(), poisoned_seen.load(),
           (unsigned long long)retried, overlapped.load(), bad.load());
    if (bad || multi == 0 || poisoned_seen == 0 || retried == 0 || overlapped == 0) return 1;
    printf("combine ok\n");
    return 0;
}
