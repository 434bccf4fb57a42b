// Stress of kbwt::CopyPool (kolmo_amd/csrc/kbwt_copypool.h): consecutive copy() calls whose lists have unequal part
// counts, from several caller threads at once, each result checked byte for byte. Built with -fsanitize=thread by
// tests/test_host_native.py; exits non-zero on a wrong byte (ThreadSanitizer reports races on its own).
#include "kbwt_copypool.h"

#include <cstdio>
#include <random>
#include <thread>
#include <vector>

int main(int argc, char **argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 200;
    kbwt::CopyPool pool(3);
    const size_t P = kbwt::CopyPool::PART;
    std::vector<std::thread> callers;
    int bad = 0;
    std::mutex bad_mu;
    for (int c = 0; c < 3; c++)
        callers.emplace_back([&, c] {
            std::mt19937_64 rng(1234 + c);
            for (int it = 0; it < iters; it++) {
                // 1..9 segments, each 1 byte .. 4 parts long: lists of very different part counts
                const int nseg = 1 + (int)(rng() % 9);
                std::vector<std::vector<uint8_t>> src(nseg), dst(nseg);
                std::vector<kbwt::CopySeg> segs;
                for (int s = 0; s < nseg; s++) {
                    const size_t n = 1 + rng() % (4 * P);
                    src[s].resize(n);
                    dst[s].assign(n, 0);
                    for (size_t k = 0; k < n; k += 4096) src[s][k] = (uint8_t)(rng() | 1);
                    src[s][n - 1] = (uint8_t)(it + s + 1);
                    segs.push_back({dst[s].data(), src[s].data(), n});
                }
                pool.copy(segs);
                for (int s = 0; s < nseg; s++)
                    if (dst[s] != src[s]) {
                        std::lock_guard<std::mutex> lk(bad_mu);
                        bad++;
                    }
            }
        });
    for (auto &t : callers) t.join();
    if (bad) {
        fprintf(stderr, "copypool: %d segments copied wrong\n", bad);
        return 1;
    }
    printf("copypool ok\n");
    return 0;
}
