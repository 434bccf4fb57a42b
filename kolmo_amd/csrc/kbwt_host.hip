// Host side of libkbwt_hip: the C-ABI entry points of include/kbwt.h, the per-device workspace, and the round loop
// of the forward suffix sort and the steps of the inverse. Validation and return codes follow
// src/algo/sais/sais.c:12027-12082 of kolmo 0.0.2.
#include "kbwt_internal.h"
#include "../../include/kbwt.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace kbwt {
namespace {

thread_local std::string g_last_error;
thread_local kbwt_stats g_stats;

struct DevError {
    int code;
};

void fail(hipError_t e, const char *what) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    g_last_error = buf;
    fprintf(stderr, "kbwt: %s\n", buf);
    throw DevError{e == hipErrorOutOfMemory ? KBWT_ERR_NOMEM : KBWT_ERR_DEVICE};
}
#define HIPCHK(x)                                      \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) fail(e_, #x);            \
    } while (0)
#define LAUNCHCHK(what) HIPCHK(hipGetLastError())

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Workspace {
    int dev = -1;
    hipStream_t stream = nullptr;
    std::mutex mu;
    uint32_t epoch = 1;
    Buf T, U, K0, K1, V0, V1, SA, ISA, L, pos[2], grp[2], status, st_sum, st_max, samples, stage;
    // small fixed device buffers
    uint32_t *counters = nullptr;      // 16 tile counters (zeroed once)
    uint32_t *counts = nullptr;        // MAXP * 256 histogram accumulators, zero between uses
    uint32_t *tot = nullptr;           // 256 byte counts
    uint32_t *dbase = nullptr;         // MAXP * 256
    uint32_t *cbase = nullptr;         // 256
    uint8_t *code = nullptr;           // 256
    unsigned long long *orand = nullptr;  // 2
    uint32_t *m_out = nullptr;         // 1
    // pinned host mirror
    uint32_t *h_counts = nullptr;      // 256
    uint8_t *h_code = nullptr;         // 256
    unsigned long long *h_orand = nullptr;  // 2 (also staging for the reset value)
    unsigned long long *h_orand_init = nullptr;
    uint32_t *h_m = nullptr;
    uint64_t *h_link = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ready = false;
    // per-kernel device time (HIP events around each launch, on the launch's stream) while profiling is on
    struct Pending { int slot; hipEvent_t a, b; uint64_t elems; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> spare;
};

struct KernelTime {
    std::string name;
    double ms = 0;
    int64_t launches = 0;
    int64_t elems = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<KernelTime> g_prof;

int prof_slot(const char *name) {
    for (size_t i = 0; i < g_prof.size(); i++) if (g_prof[i].name == name) return (int)i;
    g_prof.push_back(KernelTime{name});
    return (int)g_prof.size() - 1;
}

Workspace g_ws[64];
std::mutex g_ws_mu;

void grow(Buf &b, size_t bytes, bool zero = false) {
    if (bytes < 256) bytes = 256;
    if (b.cap >= bytes) return;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t cap = bytes + bytes / 8;   // room to grow a little without a new allocation
    HIPCHK(hipMalloc(&b.p, cap));
    if (zero) {
        HIPCHK(hipMemset(b.p, 0, cap));
        HIPCHK(hipDeviceSynchronize());
    }
    b.cap = cap;
}

Workspace &workspace() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) fail(hipErrorInvalidDevice, "hipGetDevice");
    Workspace &w = g_ws[dev];
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (!w.ready) {
        w.dev = dev;
        HIPCHK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&w.counters, 16 * sizeof(uint32_t)));
        HIPCHK(hipMemset(w.counters, 0, 16 * sizeof(uint32_t)));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMalloc(&w.counts, MAXP * 256 * sizeof(uint32_t)));
        HIPCHK(hipMemset(w.counts, 0, MAXP * 256 * sizeof(uint32_t)));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMalloc(&w.tot, 256 * sizeof(uint32_t)));
        HIPCHK(hipMalloc(&w.dbase, MAXP * 256 * sizeof(uint32_t)));
        HIPCHK(hipMalloc(&w.cbase, 256 * sizeof(uint32_t)));
        HIPCHK(hipMalloc(&w.code, 256));
        HIPCHK(hipMalloc(&w.orand, 2 * sizeof(unsigned long long)));
        HIPCHK(hipMalloc(&w.m_out, sizeof(uint32_t)));
        HIPCHK(hipHostMalloc(&w.h_counts, 256 * sizeof(uint32_t), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_code, 256, hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_orand, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_orand_init, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_m, sizeof(uint32_t), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_link, sizeof(uint64_t), hipHostMallocDefault));
        w.h_orand_init[0] = 0ull;
        w.h_orand_init[1] = ~0ull;
        HIPCHK(hipEventCreate(&w.ev0));
        HIPCHK(hipEventCreate(&w.ev1));
        w.ready = true;
    }
    return w;
}

// Launch `fn` on stream s; with profiling on, bracket it with a pair of events.
template <class F> void timed(Workspace &w, hipStream_t s, const char *name, uint64_t elems, F fn) {
    bool on;
    int slot = 0;
    {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        on = g_prof_on;
        if (on) slot = prof_slot(name);
    }
    if (!on) { fn(); return; }
    hipEvent_t a, b;
    if (w.spare.size() >= 2) { a = w.spare.back(); w.spare.pop_back(); b = w.spare.back(); w.spare.pop_back(); }
    else { HIPCHK(hipEventCreate(&a)); HIPCHK(hipEventCreate(&b)); }
    HIPCHK(hipEventRecord(a, s));
    fn();
    HIPCHK(hipEventRecord(b, s));
    w.pending.push_back({slot, a, b, elems});
}
// After a synchronisation of the stream: fold the finished launches into the per-kernel totals.
void prof_collect(Workspace &w) {
    if (w.pending.empty()) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &p : w.pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        if (p.slot < (int)g_prof.size()) {
            g_prof[p.slot].ms += ms;
            g_prof[p.slot].launches += 1;
            g_prof[p.slot].elems += (int64_t)p.elems;
        }
        w.spare.push_back(p.a);
        w.spare.push_back(p.b);
    }
    w.pending.clear();
}

inline uint32_t ceil_log2(uint32_t x) {
    uint32_t b = 0;
    while (((uint64_t)1 << b) < x) b++;
    return b;
}
template <class T> T *P(Buf &b) { return reinterpret_cast<T *>(b.p); }

uint32_t hist_grid(uint32_t n) { return std::max(1u, std::min<uint32_t>(HIST_GRID, (n + 8191) / 8192)); }

#define TIMED(name, elems, call) timed(w, s, name, (uint64_t)(elems), [&] { call; })

uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    long x = strtol(v, nullptr, 10);
    return x < (long)lo ? lo : (x > (long)hi ? hi : (uint32_t)x);
}

// Clears the histogram accumulators: they are left at zero by every scan, and this makes a call that follows an
// interrupted one start clean too.
void reset_counts(Workspace &w, hipStream_t s) { launch_fill_u8(reinterpret_cast<uint8_t *>(w.counts), 0, MAXP * 256 * 4, s); }

// ------------------------------------------------------------------ forward

void forward(Workspace &w, const uint8_t *dT, uint8_t *dU, uint32_t n, uint32_t r, int32_t *I, hipStream_t s) {
    kbwt_stats &st = g_stats;
    memset(&st, 0, sizeof st);
    st.n = (int32_t)n;
    const size_t N = std::max<size_t>(n, 64);
    grow(w.K0, N * 8);
    grow(w.K1, N * 8);
    grow(w.V0, N * 4);
    grow(w.V1, N * 4);
    grow(w.SA, N * 4);
    grow(w.ISA, N * 4);
    grow(w.L, N);
    for (int i = 0; i < 2; i++) { grow(w.pos[i], N * 4); grow(w.grp[i], N * 4); }
    const uint32_t ntiles_max = (uint32_t)((N + std::min(SORT_TILE, RG_TILE) - 1) / std::min(SORT_TILE, RG_TILE));
    grow(w.status, (size_t)ntiles_max * 256 * 8, true);
    grow(w.st_sum, (size_t)ntiles_max * 8, true);
    grow(w.st_max, (size_t)ntiles_max * 8, true);
    const uint32_t nsamp = (n - 1) / r + 1;
    grow(w.samples, (size_t)nsamp * 4);

    HIPCHK(hipEventRecord(w.ev0, s));
    reset_counts(w, s);
    // 1. byte histogram -> dense symbol code
    const uint32_t g1 = hist_grid(n);
    TIMED("byte_hist", n, launch_byte_hist(dT, n, w.counts, g1, s));
    TIMED("scan_digits", 256, launch_scan_digits(w.counts, 1, w.tot, nullptr, 0, s));
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(w.h_counts, w.tot, 256 * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint32_t sigma = 0;
    for (int c = 0; c < 256; c++) { w.h_code[c] = (uint8_t)sigma; if (w.h_counts[c]) sigma++; }
    for (int c = 0; c < 256; c++) if (!w.h_counts[c]) w.h_code[c] = 0;
    // 56 key bits above the BWT byte: 8 symbols when a symbol needs 5-7 bits, 7 bytes of binary data, and as many
    // symbols as fit for small alphabets (14 of 4 bits, 18 of 3, 28 of 2, 56 of 1)
    const uint32_t wbits = sigma <= 2 ? 1 : ceil_log2(sigma);
    const uint32_t h0 = wbits >= 5 ? std::min<uint32_t>(8, 56 / wbits) : 56 / wbits;
    const uint32_t P0 = (h0 * wbits + 7) / 8;
    st.sigma = (int32_t)sigma;
    st.bits_per_symbol = (int32_t)wbits;
    st.h0 = (int32_t)h0;
    HIPCHK(hipMemcpyAsync(w.code, w.h_code, 256, hipMemcpyHostToDevice, s));

    // 2. digit histograms of the first keys
    TIMED("hist_T", n, launch_hist_T(dT, n, w.code, wbits, h0, P0, w.counts, g1, s));
    TIMED("scan_digits", 256 * P0, launch_scan_digits(w.counts, P0, nullptr, w.dbase, 0, s));
    LAUNCHCHK();

    // 3. the first sort: P0 stable passes over the digits above the BWT byte, the last one writing the suffix array
    uint64_t *K[2] = {P<uint64_t>(w.K0), P<uint64_t>(w.K1)};
    uint32_t *V[2] = {P<uint32_t>(w.V0), P<uint32_t>(w.V1)};
    uint32_t *SA = P<uint32_t>(w.SA), *ISA = P<uint32_t>(w.ISA);
    uint8_t *L = P<uint8_t>(w.L);
    for (uint32_t p = 0; p < P0; p++) {
        SortArgs a{};
        a.from_t = p == 0;
        a.kin = p ? K[(p - 1) & 1] : nullptr;
        a.vin = p ? V[(p - 1) & 1] : nullptr;
        a.kout = K[p & 1];
        a.vout = p == P0 - 1 ? SA : V[p & 1];
        a.m = n;
        a.shift = 8 * p + 8;
        a.dbase = w.dbase + p * 256;
        a.status = P<uint64_t>(w.status);
        a.counter = w.counters + 0;
        a.epoch = w.epoch++;
        a.T = dT;
        a.code = w.code;
        a.w = wbits;
        a.h0 = h0;
        TIMED(p == 0 ? "sort_pass_T" : "sort_pass", n, launch_sort_pass(a, s));
        LAUNCHCHK();
    }
    st.passes[0] = (int32_t)P0;
    st.active[0] = n;

    // 4. groups after the first sort; the BWT byte of every row is its key's low byte
    {
        RegroupArgs g{};
        g.K = K[(P0 - 1) & 1]; g.V = SA; g.pos = nullptr; g.m = n; g.kshift = 8;
        g.SA = nullptr; g.ISA = ISA; g.L = L; g.T = nullptr;
        g.pos_out = P<uint32_t>(w.pos[0]); g.grp_out = P<uint32_t>(w.grp[0]); g.m_out = w.m_out;
        g.st_sum = P<uint64_t>(w.st_sum); g.st_max = P<uint64_t>(w.st_max); g.counter = w.counters + 1;
        g.epoch = w.epoch++;
        TIMED("regroup", n, launch_regroup(g, s));
    }
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(w.h_m, w.m_out, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint32_t m = *w.h_m;

    // 5. doubling rounds over the rows still in groups of two or more
    uint64_t h = h0;
    int cur = 0, round = 0;
    while (m > 0) {
        round++;
        if (round >= KBWT_MAX_ROUNDS) fail(hipErrorUnknown, "doubling did not converge");
        const uint32_t hh = (uint32_t)std::min<uint64_t>(h, n);
        HIPCHK(hipMemcpyAsync(w.orand, w.h_orand_init, 16, hipMemcpyHostToDevice, s));
        TIMED("round_keys", m,
              launch_round_keys(P<uint32_t>(w.pos[cur]), P<uint32_t>(w.grp[cur]), m, SA, ISA, n, hh, K[0], V[0],
                                w.orand, s));
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(w.h_orand, w.orand, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const uint64_t vary = w.h_orand[0] ^ w.h_orand[1];
        Shifts sh{};
        for (uint32_t d = 0; d < 8; d++)
            if ((vary >> (8 * d)) & 0xFFull) sh.s[sh.P++] = 8 * d;
        int src = 0;
        if (sh.P) {
            const uint32_t g2 = hist_grid(m);
            TIMED("hist_keys", m, launch_hist_keys(K[0], m, sh, w.counts, g2, s));
            TIMED("scan_digits", 256 * sh.P, launch_scan_digits(w.counts, sh.P, nullptr, w.dbase, 0, s));
            LAUNCHCHK();
            for (uint32_t p = 0; p < sh.P; p++) {
                SortArgs a{};
                a.from_t = false;
                a.kin = K[src];
                a.vin = V[src];
                a.kout = K[src ^ 1];
                a.vout = V[src ^ 1];
                a.m = m;
                a.shift = sh.s[p];
                a.dbase = w.dbase + p * 256;
                a.status = P<uint64_t>(w.status);
                a.counter = w.counters + 0;
                a.epoch = w.epoch++;
                TIMED("sort_pass", m, launch_sort_pass(a, s));
                LAUNCHCHK();
                src ^= 1;
            }
        }
        {
            RegroupArgs g{};
            g.K = K[src]; g.V = V[src]; g.pos = P<uint32_t>(w.pos[cur]); g.m = m; g.kshift = 0;
            g.SA = SA; g.ISA = ISA; g.L = L; g.T = dT;
            g.pos_out = P<uint32_t>(w.pos[cur ^ 1]); g.grp_out = P<uint32_t>(w.grp[cur ^ 1]); g.m_out = w.m_out;
            g.st_sum = P<uint64_t>(w.st_sum); g.st_max = P<uint64_t>(w.st_max); g.counter = w.counters + 1;
            g.epoch = w.epoch++;
            TIMED("regroup", m, launch_regroup(g, s));
        }
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(w.h_m, w.m_out, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        st.passes[round] = (int32_t)sh.P;
        st.active[round] = m;
        m = *w.h_m;
        cur ^= 1;
        h *= 2;
    }
    st.rounds = round;

    // 6. BWT bytes (the row bytes shifted around the primary row) and sampled indexes
    TIMED("bwt", n, launch_bwt(dT, L, ISA, dU, n, s));
    TIMED("samples", nsamp, launch_samples(ISA, r, nsamp, P<int32_t>(w.samples), s));
    LAUNCHCHK();
    HIPCHK(hipEventRecord(w.ev1, s));
    HIPCHK(hipMemcpyAsync(I, w.samples.p, (size_t)nsamp * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, w.ev0, w.ev1));
    st.ms_fwd = ms;
    prof_collect(w);
}

// ------------------------------------------------------------------ inverse

void inverse(Workspace &w, const uint8_t *dL, uint8_t *dOut, uint32_t n, uint32_t Pidx, hipStream_t s) {
    kbwt_stats &st = g_stats;
    // one splitter row in D = 2^dshift; the first cap bytes of each sublist are staged during the counting walk
    const uint32_t dshift = env_u32("KBWT_INV_DSHIFT", 6, 2, 12);
    const uint32_t cap = env_u32("KBWT_INV_CAP", 4u << dshift, 1, 1u << 16);
    const size_t N = std::max<size_t>(n, 64);
    const uint32_t dmask = (1u << dshift) - 1u;
    const uint32_t nreg = (n >> dshift) + 1;
    const uint32_t S = nreg + ((Pidx & dmask) != 0 ? 1 : 0);
    grow(w.V0, N * 4);                                 // lf
    grow(w.pos[0], ((size_t)S + 1) * 4);               // sublist lengths
    grow(w.pos[1], ((size_t)S + 1) * 4);               // continuation rows
    grow(w.K0, ((size_t)S + 1) * 8);                   // links (ping)
    grow(w.K1, ((size_t)S + 1) * 8);                   // links (pong)
    grow(w.stage, (size_t)S * cap);
    const uint32_t ntiles = (uint32_t)((N + LF_TILE - 1) / LF_TILE);
    grow(w.status, (size_t)ntiles * 256 * 8, true);

    HIPCHK(hipEventRecord(w.ev0, s));
    reset_counts(w, s);
    const uint32_t g1 = hist_grid(n);
    TIMED("byte_hist", n, launch_byte_hist(dL, n, w.counts, g1, s));
    TIMED("scan_digits", 256, launch_scan_digits(w.counts, 1, nullptr, w.cbase, 1, s));
    uint32_t *lf = P<uint32_t>(w.V0);
    TIMED("lf_pass", n, launch_lf_pass(dL, n, w.cbase, lf, P<uint64_t>(w.status), w.counters + 2, w.epoch++, s));
    LAUNCHCHK();

    st.splitters = (int32_t)S;
    uint64_t *link[2] = {P<uint64_t>(w.K0), P<uint64_t>(w.K1)};
    uint32_t *len = P<uint32_t>(w.pos[0]), *cont = P<uint32_t>(w.pos[1]);
    uint8_t *stage = P<uint8_t>(w.stage);
    TIMED("walk_count", n, launch_walk_count(lf, n, Pidx, dshift, nreg, S, w.cbase, link[0], len, stage, cap, cont, s));
    LAUNCHCHK();
    const uint32_t iters = ceil_log2(S + 1) + 1;
    int cur = 0;
    for (uint32_t it = 0; it < iters; it++) {
        TIMED("rank_jump", S, launch_rank_jump(link[cur], link[cur ^ 1], S, s));
        cur ^= 1;
    }
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(w.h_link, link[cur], 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const uint64_t l0 = *w.h_link;
    const bool valid = (uint32_t)l0 == S && (uint32_t)(l0 >> 32) == n;
    if (valid)
        TIMED("walk_emit", n, launch_walk_emit(lf, n, Pidx, S, link[cur], len, w.cbase, stage, cap, cont, dOut, s));
    else launch_fill_u8(dOut, 0, n, s);   // not a BWT: deterministic bytes that fail the caller's checksum
    LAUNCHCHK();
    HIPCHK(hipEventRecord(w.ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, w.ev0, w.ev1));
    st.ms_inv = ms;
    prof_collect(w);
}

}  // namespace
}  // namespace kbwt

using namespace kbwt;

template <class F> static int32_t guarded(F f) {
    g_last_error.clear();
    try {
        f();
        return 0;
    } catch (const DevError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return KBWT_ERR_NOMEM;
    }
}

extern "C" {

int32_t kolmo_sais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, int32_t *I,
                           int32_t threads) {
    (void)threads;
    if (T == NULL || U == NULL || A == NULL || n < 0 || r < 2 || (r & (r - 1)) != 0 || I == NULL) return -1;
    if (n <= 1) {
        if (n == 1) U[0] = T[0];
        I[0] = n;
        return 0;
    }
    return guarded([&] {
        Workspace &w = workspace();
        std::lock_guard<std::mutex> lk(w.mu);
        grow(w.T, (size_t)n + 64);
        grow(w.U, (size_t)n + 64);
        HIPCHK(hipMemcpyAsync(w.T.p, T, (size_t)n, hipMemcpyHostToDevice, w.stream));
        forward(w, P<uint8_t>(w.T), P<uint8_t>(w.U), (uint32_t)n, (uint32_t)r, I, w.stream);
        HIPCHK(hipMemcpyAsync(U, w.U.p, (size_t)n, hipMemcpyDeviceToHost, w.stream));
        HIPCHK(hipStreamSynchronize(w.stream));
    });
}

int32_t kolmo_sais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, const int32_t *I,
                             int32_t threads) {
    (void)threads;
    if (T == NULL || U == NULL || A == NULL || n < 0 || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) || I == NULL)
        return -1;
    if (n <= 1) {
        if (I[0] != n) return -1;
        if (n == 1) U[0] = T[0];
        return 0;
    }
    for (int64_t t = 0; t <= (n - 1) / r; ++t)
        if (I[t] <= 0 || I[t] > n) return -1;
    return guarded([&] {
        Workspace &w = workspace();
        std::lock_guard<std::mutex> lk(w.mu);
        grow(w.T, (size_t)n + 64);
        HIPCHK(hipMemcpyAsync(w.T.p, T, (size_t)n, hipMemcpyHostToDevice, w.stream));
        inverse(w, P<uint8_t>(w.T), P<uint8_t>(w.T), (uint32_t)n, (uint32_t)I[0], w.stream);
        HIPCHK(hipMemcpyAsync(U, w.T.p, (size_t)n, hipMemcpyDeviceToHost, w.stream));
        HIPCHK(hipStreamSynchronize(w.stream));
    });
}

int32_t kolmo_sais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t i, int32_t threads) {
    return kolmo_sais_unbwt_aux(T, U, A, n, n, &i, threads);
}

int32_t kbwt_dev_bwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, int32_t *I, void *stream) {
    if (dT == NULL || dU == NULL || n < 0 || r < 2 || (r & (r - 1)) != 0 || I == NULL) return -1;
    return guarded([&] {
        Workspace &w = workspace();
        std::lock_guard<std::mutex> lk(w.mu);
        hipStream_t s = stream ? (hipStream_t)stream : w.stream;
        if (n <= 1) {
            if (n == 1) HIPCHK(hipMemcpyAsync(dU, dT, 1, hipMemcpyDeviceToDevice, s));
            I[0] = n;
            HIPCHK(hipStreamSynchronize(s));
            return;
        }
        const uint8_t *src = dT;
        if (dT == dU || ((uintptr_t)dT & 15)) {   // in place, or not 16-byte aligned: work from a copy
            grow(w.T, (size_t)n + 64);
            HIPCHK(hipMemcpyAsync(w.T.p, dT, (size_t)n, hipMemcpyDeviceToDevice, s));
            src = P<uint8_t>(w.T);
        }
        forward(w, src, dU, (uint32_t)n, (uint32_t)r, I, s);
    });
}

int32_t kbwt_dev_unbwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, const int32_t *I, void *stream) {
    if (dT == NULL || dU == NULL || n < 0 || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) || I == NULL) return -1;
    if (n <= 1 && I[0] != n) return -1;
    if (n > 1)
        for (int64_t t = 0; t <= (n - 1) / r; ++t)
            if (I[t] <= 0 || I[t] > n) return -1;
    return guarded([&] {
        Workspace &w = workspace();
        std::lock_guard<std::mutex> lk(w.mu);
        hipStream_t s = stream ? (hipStream_t)stream : w.stream;
        if (n <= 1) {
            if (n == 1) HIPCHK(hipMemcpyAsync(dU, dT, 1, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            return;
        }
        const uint8_t *src = dT;
        if ((uintptr_t)dT & 15) {   // the byte histogram reads whole words: work from an aligned copy
            grow(w.T, (size_t)n + 64);
            HIPCHK(hipMemcpyAsync(w.T.p, dT, (size_t)n, hipMemcpyDeviceToDevice, s));
            src = P<uint8_t>(w.T);
        }
        inverse(w, src, dU, (uint32_t)n, (uint32_t)I[0], s);
    });
}

void kbwt_profile(int32_t on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
    g_prof.clear();
}

int32_t kbwt_profile_read(kbwt_kernel_time *out, int32_t max) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    int32_t k = 0;
    for (auto &t : g_prof) {
        if (k >= max) break;
        memset(&out[k], 0, sizeof out[k]);
        strncpy(out[k].name, t.name.c_str(), sizeof out[k].name - 1);
        out[k].ms = t.ms;
        out[k].launches = t.launches;
        out[k].elems = t.elems;
        k++;
    }
    return k;
}

int32_t kbwt_dev_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return -1;
    return guarded([&] {
        Workspace &w = workspace();
        hipStream_t s = stream ? (hipStream_t)stream : w.stream;
        launch_copy16(dst, src, bytes, s);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(s));
    });
}

int32_t kbwt_last_stats(kbwt_stats *out) {
    if (!out) return -1;
    *out = g_stats;
    return 0;
}

const char *kbwt_last_error(void) { return g_last_error.c_str(); }

int32_t kbwt_device_ok(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

void kbwt_release(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    Workspace &w = g_ws[dev];
    std::lock_guard<std::mutex> lk(w.mu);
    Buf *bufs[] = {&w.T, &w.U, &w.K0, &w.K1, &w.V0, &w.V1, &w.SA, &w.ISA, &w.L, &w.pos[0], &w.pos[1],
                   &w.grp[0], &w.grp[1], &w.status, &w.st_sum, &w.st_max, &w.samples, &w.stage};
    for (Buf *b : bufs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->cap = 0;
    }
}

}  // extern "C"
