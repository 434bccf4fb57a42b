// The library's CPU path: the same forward and inverse BWT on host threads, used only when a call's device work
// fails and the caller has enabled the fallback (kbwt_set_cpu_fallback, KBWT_CPU_FALLBACK=1). kolmo's Metal lane made
// the same promise: a GPU error sorts on the CPU with identical bytes (src/arch/metal/cubwt_metal.c:4-7), counted as a
// fallback (src/core/hetero.c:62-66).
//
// It restates the device algorithm, not a CPU suffix sorter:
//   forward: prefix doubling over rank groups. Suffixes are bucketed by their first two bytes and sorted by their
//            first eight (big-endian, zero-padded); every group of equal keys then sorts by (ISA[i+h] + h, or n-1-i
//            when i+h passes the end), h = 8, 16, 32, ..., until every group is a single suffix. ISA holds each
//            suffix's group head (the group's first suffix-array row). Groups are shared over threads; each round
//            gathers every key before any rank changes.
//   inverse: LF by histogram and a running count, then one walk per sampled index (chunk k decodes text
//            [k*r, (k+1)*r) backwards from the row of suffix (k+1)*r, which the stream records as I[k+1]). Every
//            chunk must end on the row its own sample names; otherwise (samples that do not match the BWT) one walk
//            from row 0 decodes the whole block, as the device path does, and a string that is no BWT decodes to
//            zeros as there.
// Output conventions are libsais': U[0] = T[n-1], the row of suffix 0 skipped, I[k] = 1 + rank(k*r)
// (src/algo/sais/sais.c:12027-12053).
#include "kbwt_cpu.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace kbwt {
namespace {

// fn(k) for k in [0, count) on up to `threads` threads (the caller takes part), k taken in order from a counter.
template <class F> void par_for(size_t count, unsigned threads, F fn) {
    if (count == 0) return;
    threads = (unsigned)std::min<size_t>(std::max(1u, threads), count);
    std::atomic<size_t> next{0};
    auto body = [&] {
        for (;;) {
            const size_t k = next.fetch_add(1, std::memory_order_relaxed);
            if (k >= count) return;
            fn(k);
        }
    };
    std::vector<std::thread> th;
    th.reserve(threads - 1);
    for (unsigned t = 1; t < threads; t++) th.emplace_back(body);
    body();
    for (auto &t : th) t.join();
}

inline uint64_t key8(const uint8_t *T, uint32_t n, uint32_t i) {
    uint64_t k = 0;
    if ((uint64_t)i + 8 <= n) {
        for (int j = 0; j < 8; j++) k = (k << 8) | T[i + j];
        return k;
    }
    for (uint32_t j = 0; j < 8; j++) k = (k << 8) | (i + j < n ? T[i + j] : 0u);
    return k;
}

struct Group {
    uint32_t head, size;
};

// SA and ISA (group heads) of T[0..n), n >= 2.
void suffix_sort(const uint8_t *T, uint32_t n, unsigned threads, std::vector<uint32_t> &SA, std::vector<uint32_t> &ISA) {
    SA.resize(n);
    ISA.resize(n);
    // 1. buckets of the first two bytes (a suffix of one byte has second byte 0 and sorts later by its key)
    constexpr uint32_t NB = 65536;
    std::vector<uint32_t> bstart(NB + 1, 0);
    auto b2 = [&](uint32_t i) { return ((uint32_t)T[i] << 8) | (i + 1 < n ? T[i + 1] : 0u); };
    for (uint32_t i = 0; i < n; i++) bstart[b2(i) + 1]++;
    for (uint32_t b = 0; b < NB; b++) bstart[b + 1] += bstart[b];
    {
        std::vector<uint32_t> at(bstart.begin(), bstart.end() - 1);
        for (uint32_t i = 0; i < n; i++) SA[at[b2(i)]++] = i;
    }
    // 2. each bucket by the first eight bytes; groups of equal keys
    std::vector<std::vector<Group>> found(NB);
    par_for(NB, threads, [&](size_t b) {
        const uint32_t lo = bstart[b], hi = bstart[b + 1];
        if (hi - lo > 1)
            std::sort(SA.begin() + lo, SA.begin() + hi,
                      [&](uint32_t x, uint32_t y) { return key8(T, n, x) < key8(T, n, y); });
        uint32_t head = lo;
        uint64_t prev = hi > lo ? key8(T, n, SA[lo]) : 0;
        for (uint32_t p = lo; p < hi; p++) {
            const uint64_t k = key8(T, n, SA[p]);
            if (k != prev) {
                if (p - head > 1) found[b].push_back({head, p - head});
                head = p;
                prev = k;
            }
            ISA[SA[p]] = head;
        }
        if (hi - head > 1) found[b].push_back({head, hi - head});
    });
    std::vector<Group> groups;
    for (auto &f : found) {
        groups.insert(groups.end(), f.begin(), f.end());
        std::vector<Group>().swap(f);
    }
    // 3. doubling rounds over the groups of two or more suffixes
    std::vector<uint64_t> key(n);   // (sec << 32 | suffix) of every row of a group, in SA order
    for (uint64_t h = 8; !groups.empty(); h *= 2) {
        // gather every key of the round before any rank changes
        par_for(groups.size(), threads, [&](size_t g) {
            const Group gr = groups[g];
            for (uint32_t p = gr.head; p < gr.head + gr.size; p++) {
                const uint32_t i = SA[p];
                const uint64_t sec = (uint64_t)i + h < n ? (uint64_t)ISA[i + h] + h : (uint64_t)(n - 1 - i);
                key[p] = (sec << 32) | i;
            }
        });
        std::vector<std::vector<Group>> next(std::max<size_t>(1, std::min<size_t>(groups.size(), 4096)));
        const size_t per = (groups.size() + next.size() - 1) / next.size();
        par_for(next.size(), threads, [&](size_t c) {
            for (size_t g = c * per; g < std::min(groups.size(), (c + 1) * per); g++) {
                const Group gr = groups[g];
                uint64_t *k = key.data() + gr.head;
                std::sort(k, k + gr.size);
                uint32_t head = gr.head;
                for (uint32_t j = 0; j < gr.size; j++) {
                    const uint32_t p = gr.head + j;
                    if (j > 0 && (k[j] >> 32) != (k[j - 1] >> 32)) {
                        if (p - head > 1) next[c].push_back({head, p - head});
                        head = p;
                    }
                    SA[p] = (uint32_t)k[j];
                    ISA[(uint32_t)k[j]] = head;
                }
                if (gr.head + gr.size - head > 1) next[c].push_back({head, gr.head + gr.size - head});
            }
        });
        groups.clear();
        for (auto &f : next) groups.insert(groups.end(), f.begin(), f.end());
    }
}

}  // namespace

int32_t cpu_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I, int32_t threads) {
    if (n <= 1) {
        if (n == 1) U[0] = T[0];
        I[0] = n;
        return 0;
    }
    try {
        const unsigned th = (unsigned)std::max(1, std::min<int32_t>(threads, 64));
        std::vector<uint32_t> SA, ISA;
        suffix_sort(T, (uint32_t)n, th, SA, ISA);
        for (int32_t k = 0; k <= (n - 1) / r; k++) I[k] = 1 + (int32_t)ISA[(size_t)k * r];
        std::vector<uint8_t> L((size_t)n);
        L[0] = T[n - 1];
        size_t q = 1;
        for (int32_t p = 0; p < n; p++)
            if (SA[p] != 0) L[q++] = T[SA[p] - 1];
        memcpy(U, L.data(), (size_t)n);   // T == U is allowed
        return 0;
    } catch (const std::bad_alloc &) {
        return -2;
    }
}

int32_t cpu_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, const int32_t *I, int32_t threads) {
    if (n <= 1) {
        if (n == 1) U[0] = T[0];
        return 0;
    }
    try {
        const uint32_t N = (uint32_t)n, P = (uint32_t)I[0];
        // rows 0..n of the matrix: row 0 is the end marker's suffix, row P holds suffix 0 (no byte stored); stored
        // byte q is row q before P and row q + 1 from P on
        std::vector<uint32_t> C(257, 0);
        for (uint32_t q = 0; q < N; q++) C[T[q] + 1]++;
        C[0] = 1;
        for (int c = 0; c < 256; c++) C[c + 1] += C[c];
        std::vector<uint32_t> LF((size_t)N + 1, 0);
        std::vector<uint8_t> Lr((size_t)N + 1, 0);
        {
            std::vector<uint32_t> occ(C.begin(), C.end() - 1);
            for (uint32_t y = 0; y <= N; y++) {
                if (y == P) continue;
                const uint8_t c = T[y < P ? y : y - 1];
                Lr[y] = c;
                LF[y] = occ[c]++;
            }
        }
        std::vector<uint8_t> out((size_t)N);
        // one chunk per sampled index, each checked against the sample that names its last row
        const uint32_t R = (uint32_t)r, K = (N - 1) / R + 1;
        std::atomic<bool> ok{true};
        par_for(K, (unsigned)std::max(1, std::min<int32_t>(threads, 64)), [&](size_t k) {
            const uint32_t lo = (uint32_t)k * R, hi = std::min<uint64_t>((uint64_t)(k + 1) * R, N);
            uint32_t y = hi == N ? 0u : (uint32_t)I[k + 1];
            for (uint32_t e = hi; e > lo; e--) {
                if (y == P) { ok = false; return; }
                out[e - 1] = Lr[y];
                y = LF[y];
            }
            if (y != (uint32_t)I[k]) ok = false;
        });
        if (!ok) {   // the samples do not match this BWT: one walk from the end marker's row, as the device path does
            uint32_t y = 0;
            bool valid = true;
            for (uint32_t e = N; e > 0; e--) {
                if (y == P) { valid = false; break; }
                out[e - 1] = Lr[y];
                y = LF[y];
            }
            if (!valid || y != P) std::fill(out.begin(), out.end(), 0);
        }
        memcpy(U, out.data(), (size_t)N);
        return 0;
    } catch (const std::bad_alloc &) {
        return -2;
    }
}

}  // namespace kbwt
