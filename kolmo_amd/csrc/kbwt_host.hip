// Host side of libkbwt_hip: the C-ABI entry points of include/kbwt.h, the per-device workspace, and the steps of the
// forward suffix sort (round loop) and of the inverse, for a batch of blocks (one block is the batch of one);
// concurrent kolmo_sais_* calls are gathered into batches; host buffers move through pinned double buffers; the
// per-size-class tuning table and the autotuner.
// Validation and return codes follow src/algo/sais/sais.c:12027-12082 of kolmo 0.0.2.
#include "kbwt_internal.h"
#include "kbwt_combine.h"
#include "kbwt_copypool.h"
#include "kbwt_cpu.h"
#include "kbwt_router.h"
#include "../../include/kbwt.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace kbwt {
namespace {

constexpr size_t STAGE_MAX = (size_t)8 << 30;   // inverse: bytes of the staged sublist heads at most

thread_local std::string g_last_error;
thread_local kbwt_stats g_stats;

struct DevError {
    int code;
};

[[noreturn]] void fail(hipError_t e, const char *what) {
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

// ---- counters of the error path (kbwt_get_counters), and host-side fault injection
//
// A fault point is a place where the runtime can report an error: a device allocation, a kernel launch, a stream
// synchronisation, a host <-> device copy. kbwt_fault_inject("site:k[:count]") (or KBWT_FAULT in the environment at
// load) makes occurrences k .. k+count-1 of that site, counted from the call that armed it, fail before the runtime is
// called, as if it had returned an error. Nothing is launched differently: every kernel queued before the fault
// completes normally, and the error path (drain the stream, reset the cached state, run failed blocks again on their
// own, the optional CPU path) is exercised without a fault on the device.
enum FaultSite { FS_ALLOC, FS_LAUNCH, FS_SYNC, FS_COPY, FS_N };
const char *const FAULT_NAMES[FS_N] = {"alloc", "launch", "sync", "copy"};
struct FaultPlan {
    std::atomic<int> site{-1};
    std::atomic<uint64_t> at{0}, count{0}, seen{0};
};
FaultPlan g_fault;
struct Counters {
    std::atomic<int64_t> device_errors{0}, retried_blocks{0}, cpu_fallbacks{0}, injected_faults{0}, graph_replays{0};
};
Counters g_counters;
std::atomic<int> g_cpu_fallback{-1};   // -1: not read yet from KBWT_CPU_FALLBACK

bool parse_fault(const char *spec, int &site, uint64_t &at, uint64_t &count) {
    site = -1;
    if (!spec || !*spec) return true;   // disarm
    const char *colon = strchr(spec, ':');
    if (!colon) return false;
    for (int k = 0; k < FS_N; k++)
        if ((size_t)(colon - spec) == strlen(FAULT_NAMES[k]) && strncmp(spec, FAULT_NAMES[k], colon - spec) == 0)
            site = k;
    if (site < 0) return false;
    char *end = nullptr;
    at = strtoull(colon + 1, &end, 10);
    count = 1;
    if (*end == ':') count = strtoull(end + 1, &end, 10);
    if (*end != 0 || at == 0 || count == 0) {
        site = -1;
        return false;
    }
    return true;
}
void arm_fault(int site, uint64_t at, uint64_t count) {
    g_fault.site.store(-1);
    g_fault.at.store(at);
    g_fault.count.store(count);
    g_fault.seen.store(0);
    g_fault.site.store(site);
}
struct FaultFromEnv {
    FaultFromEnv() {
        int site;
        uint64_t at, count;
        if (parse_fault(getenv("KBWT_FAULT"), site, at, count) && site >= 0) arm_fault(site, at, count);
    }
} g_fault_env;

void fault_point(FaultSite site) {
    if (g_fault.site.load(std::memory_order_relaxed) != (int)site) return;
    const uint64_t k = g_fault.seen.fetch_add(1) + 1, at = g_fault.at.load();
    if (k < at || k - at >= g_fault.count.load()) return;
    g_counters.injected_faults++;
    char what[64];
    snprintf(what, sizeof what, "%s (injected fault %llu)", FAULT_NAMES[site], (unsigned long long)k);
    fail(site == FS_ALLOC ? hipErrorOutOfMemory : site == FS_LAUNCH ? hipErrorLaunchFailure : hipErrorUnknown, what);
}
#define LAUNCHCHK()                        \
    do {                                   \
        fault_point(FS_LAUNCH);            \
        HIPCHK(hipGetLastError());         \
    } while (0)
#define SYNCCHK(st)                              \
    do {                                         \
        fault_point(FS_SYNC);                    \
        HIPCHK(hipStreamSynchronize(st));        \
    } while (0)

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

constexpr size_t LIST_PEEK = 1024;   // large-group list entries read back with the list counters

// A pinned double buffer for host <-> device copies: the DMA of one chunk runs while CPU threads fill or drain the
// other; ev[b] marks the last DMA that used buf[b].
struct Pins {
    uint8_t *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

// One combined batch of the host-pointer entry points in flight: its packed bytes on the device, its own pinned chunks
// and copy stream, so that its host copies and DMA run while another batch's kernels do. staged: the inputs are on
// the device; computed: the kernels are done.
struct Lane {
    std::mutex mu;
    bool ready = false;
    Buf T, U;
    Pins pins;
    hipStream_t cs = nullptr;
    hipEvent_t staged = nullptr, computed = nullptr;
};

// Executable graphs of one launch sequence, by key (shapes and buffers). A key met on one of the last two calls is
// captured when it comes again, and the two graphs used last are kept, so that calls alternating between two shapes
// or two buffers (the two lanes of combined batches) both replay.
struct GraphCache {
    struct Entry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    Entry e[2];
    std::vector<uint64_t> seen[2];
    uint64_t tick = 0;
    void clear() {
        for (Entry &x : e) {
            if (x.exec) (void)hipGraphExecDestroy(x.exec);
            x.exec = nullptr;
            x.key.clear();
            x.used = 0;
        }
        seen[0].clear();
        seen[1].clear();
    }
};

struct Workspace {
    int dev = -1;
    hipStream_t stream = nullptr;
    std::mutex mu;
    uint32_t epoch = 1;
    uint32_t dev_epoch_calls = 0;           // inverse calls since the device epoch last started a cycle
    Buf T, U, K0, K1, V0, V1, SA, ISA, L, pos[2], grp[2], status, st_sum, carry, samples, stage, blk, counts, dbase,
        tot, cbase, bad, meta, gh[2], gs[2], lh[2], ls[2], go, sec, sufl, lo, tstart, act, rfeat, bigram, rmeta, eh, es,
        ecnt, eoff, pair, tmax, lrow;
    uint32_t *counters = nullptr;           // 16 tile counters (zeroed once); [15]: the device epoch of graphs
    uint32_t *gcnt = nullptr;               // 4 group-list counters
    uint32_t *h_gcnt = nullptr;
    unsigned long long *orand = nullptr;    // 2
    uint32_t *m_out = nullptr;              // 1
    // the inverse's launch sequence, and the forward's first sort, as executable graphs, for calls that repeat a
    // recent call's shape
    GraphCache inv_graph, fwd_graph;
    // pinned host staging: the host-pointer entry points' double buffer (PIN_CHUNK bytes each)
    Pins pins;
    // the combined batches of concurrent calls, one lane each (kbwt_combine.h); a lane's mutex is taken before mu
    Lane lane[COMBINE_LANES];
    void *h_meta = nullptr;
    size_t h_meta_cap = 0;
    unsigned long long *h_orand = nullptr, *h_orand_init = nullptr;
    uint32_t *h_m = nullptr;
    uint32_t *h_peek = nullptr;             // 2 * LIST_PEEK: the first heads and sizes of a large-group list
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ready = false;
    // per-kernel device time (HIP events around each launch, on the launch's stream) while profiling is on
    struct Pending {
        int slot;
        hipEvent_t a, b;
        uint64_t elems;
    };
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
std::string g_prof_only;   // when not empty, only launches of this name are timed
std::vector<KernelTime> g_prof;

bool prof_wants(const char *name) { return g_prof_on && (g_prof_only.empty() || g_prof_only == name); }

int prof_slot(const char *name) {
    for (size_t i = 0; i < g_prof.size(); i++)
        if (g_prof[i].name == name) return (int)i;
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
    fault_point(FS_ALLOC);
    HIPCHK(hipMalloc(&b.p, cap));
    if (zero) {
        HIPCHK(hipMemset(b.p, 0, cap));
        HIPCHK(hipDeviceSynchronize());
    }
    b.cap = cap;
}

void clear_status(Workspace &w, hipStream_t s) {
    if (w.status.p) launch_fill_u8(reinterpret_cast<uint8_t *>(w.status.p), 0, w.status.cap, s);
    if (w.st_sum.p) launch_fill_u8(reinterpret_cast<uint8_t *>(w.st_sum.p), 0, w.st_sum.cap, s);
}

void *host_staging(Workspace &w, size_t bytes) {
    if (w.h_meta_cap < bytes) {
        if (w.h_meta) HIPCHK(hipHostFree(w.h_meta));
        w.h_meta = nullptr;
        w.h_meta_cap = 0;
        size_t cap = bytes + bytes / 4 + 4096;
        HIPCHK(hipHostMalloc(&w.h_meta, cap, hipHostMallocDefault));
        w.h_meta_cap = cap;
    }
    return w.h_meta;
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
        HIPCHK(hipMalloc(&w.orand, 2 * sizeof(unsigned long long)));
        HIPCHK(hipMalloc(&w.m_out, sizeof(uint32_t)));
        HIPCHK(hipMalloc(&w.gcnt, 4 * sizeof(uint32_t)));
        HIPCHK(hipHostMalloc(&w.h_gcnt, 4 * sizeof(uint32_t), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_orand, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_orand_init, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_m, sizeof(uint32_t), hipHostMallocDefault));
        HIPCHK(hipHostMalloc(&w.h_peek, 2 * LIST_PEEK * sizeof(uint32_t), hipHostMallocDefault));
        w.h_orand_init[0] = 0ull;
        w.h_orand_init[1] = ~0ull;
        {   // the device epoch starts in the upper half of the epoch space
            const uint32_t e0 = EPOCH_HALF;
            HIPCHK(hipMemcpy(w.counters + 15, &e0, 4, hipMemcpyHostToDevice));
        }
        // Events with the default system-scope fence. A call records ev0 right after copying its input and metadata
        // to the device, and that record's fence is what invalidates the XCDs' L2 caches before the first kernel
        // reads the copied bytes. Without it, lines of the same buffer cached by an earlier call (the previous
        // call's decoded output, when the input buffer comes back at the same address) were read in place of the
        // new input now and then, and a block of the batch came out wrong.
        HIPCHK(hipEventCreate(&w.ev0));
        HIPCHK(hipEventCreate(&w.ev1));
        w.ready = true;
    }
    return w;
}

// ---- host buffers to and from the device through pinned memory
//
// The runtime moves pageable host memory through its own staging with one CPU thread. The host-pointer entry points
// instead move their host spans (the blocks of a call or of a combined batch, back to back on the device) through two
// pinned chunks: a small pool of CPU threads copies one chunk while the DMA engine moves the other.
constexpr size_t PIN_CHUNK = (size_t)8 << 20;
constexpr size_t PIN_MIN = (size_t)1 << 20;   // below this many bytes in all the runtime's own path is as fast

struct Span {   // host bytes
    uint8_t *p;
    size_t n;
};
CopyPool g_copy;

void release_pins(Pins &pp) {
    for (int b = 0; b < 2; b++) {
        if (pp.ev[b]) {
            (void)hipEventSynchronize(pp.ev[b]);
            (void)hipEventDestroy(pp.ev[b]);
        }
        if (pp.buf[b]) (void)hipHostFree(pp.buf[b]);
        pp.buf[b] = nullptr;
        pp.ev[b] = nullptr;
    }
}

void pin_ready(Pins &pp) {
    for (int b = 0; b < 2; b++) {
        if (!pp.buf[b]) HIPCHK(hipHostMalloc((void **)&pp.buf[b], PIN_CHUNK, hipHostMallocDefault));
        if (!pp.ev[b]) HIPCHK(hipEventCreateWithFlags(&pp.ev[b], hipEventDisableTiming));
    }
}

// The copies between the host spans, laid back to back, and pinned buffer `pin`, for bytes [o, o + len) of them.
// `at` walks the spans (span index, offset in it) and moves on by len.
std::vector<CopySeg> span_segs(const std::vector<Span> &sp, size_t len, uint8_t *pin, bool to_pin,
                               std::pair<size_t, size_t> &at) {
    std::vector<CopySeg> segs;
    for (size_t done = 0; done < len;) {
        const Span &x = sp[at.first];
        const size_t take = std::min(len - done, x.n - at.second);
        if (to_pin) segs.push_back({pin + done, x.p + at.second, take});
        else segs.push_back({x.p + at.second, pin + done, take});
        done += take;
        at.second += take;
        if (at.second == x.n) {
            at.first++;
            at.second = 0;
        }
    }
    return segs;
}

size_t span_total(const std::vector<Span> &sp) {
    size_t n = 0;
    for (const Span &x : sp) n += x.n;
    return n;
}

// Host spans to dst on stream s (returns once the last chunk is queued; the spans may then be reused).
void to_device(Pins &pp, uint8_t *dst, const std::vector<Span> &sp, hipStream_t s) {
    fault_point(FS_COPY);
    const size_t n = span_total(sp);
    if (n < PIN_MIN) {
        for (const Span &x : sp) {
            HIPCHK(hipMemcpyAsync(dst, x.p, x.n, hipMemcpyHostToDevice, s));
            dst += x.n;
        }
        return;
    }
    pin_ready(pp);
    std::pair<size_t, size_t> at{0, 0};
    while (at.first < sp.size() && sp[at.first].n == 0) at.first++;
    for (size_t o = 0, k = 0; o < n; o += PIN_CHUNK, k++) {
        const size_t len = std::min(PIN_CHUNK, n - o);
        const int b = (int)(k & 1);
        HIPCHK(hipEventSynchronize(pp.ev[b]));   // the DMA that read this buffer last is done
        g_copy.copy(span_segs(sp, len, pp.buf[b], true, at));
        HIPCHK(hipMemcpyAsync(dst + o, pp.buf[b], len, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(pp.ev[b], s));
    }
}

// src to the host spans on stream s, after the work queued on s so far (returns with the bytes in the spans).
void to_host(Pins &pp, const std::vector<Span> &sp, const uint8_t *src, hipStream_t s) {
    fault_point(FS_COPY);
    const size_t n = span_total(sp);
    if (n < PIN_MIN) {
        for (const Span &x : sp) {
            HIPCHK(hipMemcpyAsync(x.p, src, x.n, hipMemcpyDeviceToHost, s));
            src += x.n;
        }
        SYNCCHK(s);
        return;
    }
    pin_ready(pp);
    const size_t chunks = (n + PIN_CHUNK - 1) / PIN_CHUNK;
    auto issue = [&](size_t k) {
        const int b = (int)(k & 1);
        HIPCHK(hipEventSynchronize(pp.ev[b]));
        HIPCHK(hipMemcpyAsync(pp.buf[b], src + k * PIN_CHUNK, std::min(PIN_CHUNK, n - k * PIN_CHUNK),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(pp.ev[b], s));
    };
    std::pair<size_t, size_t> at{0, 0};
    while (at.first < sp.size() && sp[at.first].n == 0) at.first++;
    issue(0);
    for (size_t k = 0; k < chunks; k++) {
        if (k + 1 < chunks) issue(k + 1);   // the other buffer: its chunk was copied out in the step before
        const int b = (int)(k & 1);
        HIPCHK(hipEventSynchronize(pp.ev[b]));
        g_copy.copy(span_segs(sp, std::min(PIN_CHUNK, n - k * PIN_CHUNK), pp.buf[b], false, at));
    }
}

uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi);

// Lookback status words are never cleared between launches: an epoch tag tells a word of this launch from older
// ones. Host-side epochs stay in [1, EPOCH_HALF - 1); before they wrap, the status arrays are cleared (a fill kernel,
// ordered on the stream like the launches), so a word left by a launch of the previous cycle cannot read as ready.
void clear_status(Workspace &w, hipStream_t s);
inline uint32_t next_epoch(Workspace &w, hipStream_t s) {
    if (w.epoch >= EPOCH_HALF - 1) {
        clear_status(w, s);
        w.epoch = 1;
    }
    return w.epoch++;
}

// Whether any of the named launches would be bracketed by profiling events.
bool prof_touches(std::initializer_list<const char *> names) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (const char *n : names)
        if (prof_wants(n)) return true;
    return false;
}

// Run `body` (a fixed launch sequence on stream s, whose launches are `names`) directly, or from an executable
// graph of g: a key met on one of the last two calls is captured when it comes again, and replayed from then on.
// Graphs are skipped while profiling times one of the launches (events bracket them), on the legacy default stream
// (it cannot be captured), for combined batches (GraphsOff) and with KBWT_GRAPHS=0.
// GraphsOff: while one lives on a thread, that thread's launch sequences run directly (no capture, no replay).
thread_local int g_graphs_off = 0;
struct GraphsOff {
    GraphsOff() { g_graphs_off++; }
    ~GraphsOff() { g_graphs_off--; }
};

template <class F>
void run_graph(GraphCache &g, const std::vector<uint64_t> &key, hipStream_t s,
               std::initializer_list<const char *> names, F body) {
    const bool eligible =
        s != nullptr && !g_graphs_off && !prof_touches(names) && env_u32("KBWT_GRAPHS", 1, 0, 1);
    if (!eligible) {
        body();
        return;
    }
    g.tick++;
    for (GraphCache::Entry &x : g.e)
        if (x.exec && x.key == key) {
            x.used = g.tick;
            HIPCHK(hipGraphLaunch(x.exec, s));
            g_counters.graph_replays++;
            return;
        }
    if (key != g.seen[0] && key != g.seen[1]) {
        g.seen[1].swap(g.seen[0]);
        g.seen[0] = key;
        body();
        return;
    }
    GraphCache::Entry &x = g.e[0].used <= g.e[1].used ? g.e[0] : g.e[1];
    if (x.exec) (void)hipGraphExecDestroy(x.exec);
    x.exec = nullptr;
    x.key.clear();
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {   // leave the stream out of capture mode before the error goes on
        if (hipStreamEndCapture(s, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&x.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    x.key = key;
    x.used = g.tick;
    HIPCHK(hipGraphLaunch(x.exec, s));
    g_counters.graph_replays++;
}

// Launch `fn` on stream s; with profiling on, bracket it with a pair of events.
template <class F> void timed(Workspace &w, hipStream_t s, const char *name, uint64_t elems, F fn) {
    bool on;
    int slot = 0;
    {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        on = prof_wants(name);
        if (on) slot = prof_slot(name);
    }
    if (!on) {
        fn();
        return;
    }
    hipEvent_t a, b;
    if (w.spare.size() >= 2) {
        a = w.spare.back();
        w.spare.pop_back();
        b = w.spare.back();
        w.spare.pop_back();
    } else {   // default fences: a pair may be recorded right after a host-to-device copy (see ev0)
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
    }
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
#define TIMED(name, elems, call) timed(w, s, name, (uint64_t)(elems), [&] { call; })

inline uint32_t ceil_log2(uint32_t x) {
    uint32_t b = 0;
    while (((uint64_t)1 << b) < x) b++;
    return b;
}
template <class T> T *P(Buf &b) { return reinterpret_cast<T *>(b.p); }

uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    long x = strtol(v, nullptr, 10);
    return x < (long)lo ? lo : (x > (long)hi ? hi : (uint32_t)x);
}

// Launch parameters per size class (ceil(log2 N) of a call's packed length), set by kbwt_autotune or
// kbwt_set_tuning. An environment variable (KBWT_SORT_CFG, KBWT_INV_DSHIFT, KBWT_SEG_RANK) overrides the table.
struct Tuning {
    uint32_t sort_cfg = 8, inv_dshift = 6, seg_rank = SEG_RANK_MAX;
};
constexpr int TUNE_CLASSES = 33;
// The built-in table, measured on one MI355X over single text + CSV blocks of 1-32 MiB: up to 2 MiB the radix pass
// takes 8 K-key tiles (more workgroups than CUs) and the inverse walks shorter sublists (splitters every 8 or 16
// rows: the longest sublist bounds the walk); to 8 MiB splitters every 16 rows; from 16 MiB on 16 K-key tiles and
// splitters every 64 rows; above 64 MiB every 128 rows (the sublists' links then stay within the caches: on a 128 MiB
// block the ranking takes 0.32 ms instead of 0.83 for a walk 0.16 ms longer).
Tuning default_tuning(uint32_t cls) {
    Tuning t;
    if (cls <= 20) {
        t.sort_cfg = 0;
        t.inv_dshift = 3;
    } else if (cls == 21) {
        t.sort_cfg = 0;
        t.inv_dshift = 4;
    } else if (cls <= 23) {
        t.inv_dshift = 4;
    } else if (cls >= 27) {
        t.inv_dshift = 7;
    }
    return t;
}
struct TuneTable {
    Tuning t[TUNE_CLASSES];
    TuneTable() {
        for (int c = 0; c < TUNE_CLASSES; c++) t[c] = default_tuning((uint32_t)c);
    }
};
std::mutex g_tune_mu;
TuneTable g_tune_table;
Tuning (&g_tune)[TUNE_CLASSES] = g_tune_table.t;
uint32_t size_class(uint32_t N) {
    uint32_t c = 0;
    while (c < 32 && ((uint64_t)1 << c) < N) c++;
    return c;
}
Tuning tuning_for_class(uint32_t cls) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune[cls];
}
Tuning tuning_for(uint32_t N) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    Tuning t = g_tune[size_class(N)];
    t.sort_cfg = env_u32("KBWT_SORT_CFG", t.sort_cfg, 0, SORT_NCFG - 1);
    t.inv_dshift = env_u32("KBWT_INV_DSHIFT", t.inv_dshift, 2, 12);
    t.seg_rank = env_u32("KBWT_SEG_RANK", t.seg_rank, 0, 1024);
    return t;
}

// first[b] = tiles of the blocks before b, first[B] = all tiles.
template <class F> uint32_t fill_tiles(uint32_t *first, uint32_t B, F tiles_of) {
    uint32_t t = 0;
    for (uint32_t b = 0; b < B; b++) {
        first[b] = t;
        t += tiles_of(b);
    }
    first[B] = t;
    return t;
}
inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// ------------------------------------------------------------------ forward

// The rest of the router's features after the doubling rounds: m_r / n from the active counts, and the order-1
// entropy from bigram histograms (blocks in groups of at most 1024, 256 KiB of histogram each).
void router_features_tail(Workspace &w, const uint8_t *dT, const BlockInfo *d_bl, const std::vector<BlockInfo> &bl,
                          const uint32_t *d_act, float *feat, hipStream_t s) {
    const uint32_t B = (uint32_t)bl.size();
    constexpr uint32_t CH = 1024;
    grow(w.rfeat, (size_t)B * ROUTER_NF * 4);
    grow(w.bigram, (size_t)std::min(B, CH) * 65536 * 4, true);   // every use clears what it filled
    grow(w.rmeta, ((size_t)CH + 1) * 4);
    std::vector<uint32_t> first(CH + 1);
    for (uint32_t b0 = 0; b0 < B; b0 += CH) {
        const uint32_t nb = std::min(CH, B - b0);
        first[0] = 0;
        for (uint32_t k = 0; k < nb; k++) first[k + 1] = first[k] + (bl[b0 + k].n + BIGRAM_TILE - 1) / BIGRAM_TILE;
        HIPCHK(hipMemcpyAsync(w.rmeta.p, first.data(), ((size_t)nb + 1) * 4, hipMemcpyHostToDevice, s));
        const Tiling tl{P<uint32_t>(w.rmeta), d_bl + b0, nb};
        launch_order1(dT, tl, first[nb], d_bl + b0, nb, P<uint32_t>(w.bigram), P<float>(w.rfeat) + (size_t)b0 * ROUTER_NF,
                      s);
        LAUNCHCHK();
        SYNCCHK(s);   // rmeta and first are reused by the next group
    }
    std::vector<float> ent((size_t)B * ROUTER_NF);
    std::vector<uint32_t> act((size_t)ROUTER_ROUNDS * B);
    HIPCHK(hipMemcpyAsync(ent.data(), w.rfeat.p, ent.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(act.data(), d_act, act.size() * 4, hipMemcpyDeviceToHost, s));
    SYNCCHK(s);
    for (uint32_t b = 0; b < B; b++) {
        float *f = feat + (size_t)b * ROUTER_NF;
        f[256] = ent[(size_t)b * ROUTER_NF + 256];
        for (int r = 0; r < ROUTER_ROUNDS; r++) f[257 + r] = (float)act[(size_t)r * B + b] / (float)bl[b].n;
    }
}

// A radix pass, after checking on the host that its grid fits what the kernel assumes: the tiles were counted with the
// configuration's own tile size, and a flat pass (no per-block tiling) has exactly ceil(m / tile) of them, so that no
// tile starts past the keys.
void sort_pass_checked(const SortArgs &a, uint32_t ntiles, uint32_t tile, hipStream_t s) {
    if (tile != sort_tile(a.cfg) ||
        (!a.tiling.first && (uint64_t)ntiles != ((uint64_t)a.m + tile - 1) / tile))
        fail(hipErrorInvalidValue, "radix pass launch does not match its configuration's tile");
    launch_sort_pass(a, ntiles, s);
}

// bl: one BlockInfo per block with off, n (>= 2), r and soff filled in; packed input dT (N bytes), output dU.
// I: host array of nsamp sampled indexes. Per-call metadata (blocks, tile tables, code tables) goes to the device
// in one copy.
// feat (host, B x ROUTER_NF, or nullptr): the router's features of every block, computed on the way.
void forward(Workspace &w, const uint8_t *dT, uint8_t *dU, std::vector<BlockInfo> &bl, uint32_t N, int32_t *I,
             uint32_t nsamp, hipStream_t s, float *feat = nullptr) {
    const uint32_t B = (uint32_t)bl.size();
    kbwt_stats &st = g_stats;
    memset(&st, 0, sizeof st);
    st.n = (int32_t)N;
    st.blocks = (int32_t)B;
    const Tuning tune = tuning_for(N);
    const int scfg = (int)tune.sort_cfg;
    const uint32_t stile = sort_tile(scfg);
    const size_t NN = std::max<size_t>(N, 64);
    grow(w.K0, NN * 8);
    grow(w.K1, NN * 8);
    grow(w.V0, NN * 4);
    grow(w.V1, NN * 4);
    grow(w.SA, NN * 4);
    grow(w.ISA, NN * 4);
    grow(w.L, NN);
    grow(w.pos[0], NN * 4);                    // rows of the large groups (SA position, group head)
    grow(w.grp[0], NN * 4);
    grow(w.carry, ((NN + RG_TILE - 1) / RG_TILE) * 4);
    // group lists of the doubling rounds (groups have two rows or more), ping-pong between rounds
    const size_t GMAX = NN / 2 + 2;
    for (int i = 0; i < 2; i++) {
        grow(w.gh[i], GMAX * 4);
        grow(w.gs[i], GMAX * 4);
        grow(w.lh[i], GMAX * 4);
        grow(w.ls[i], GMAX * 4);
    }
    const size_t ntiles_max = (NN + MIN_TILE - 1) / MIN_TILE + B;
    grow(w.status, ntiles_max * 256 * 8, true);
    grow(w.st_sum, ntiles_max * 8, true);
    grow(w.samples, (size_t)nsamp * 4);
    grow(w.counts, (size_t)B * MAXP * 256 * 4, true);
    grow(w.tot, (size_t)B * 256 * 4);
    grow(w.dbase, (size_t)B * MAXP * 256 * 4);
    if (B > 1) grow(w.blk, NN * 2);

    // metadata layout: blocks | hist tiles | sort tiles | partition tiles | codes | m-gram units; the byte counts come
    // back behind it
    const size_t o_ht = align16(B * sizeof(BlockInfo)), o_st = o_ht + (B + 1) * 4, o_pt = o_st + (B + 1) * 4;
    const size_t o_codes = align16(o_pt + (B + 1) * 4);
    const size_t o_gu = align16(o_codes + (size_t)B * 256);
    const size_t bytes = o_gu + (B + 1) * 4;
    uint8_t *hm = (uint8_t *)host_staging(w, bytes + (size_t)B * 256 * 4);
    grow(w.meta, bytes);
    uint8_t *dm = P<uint8_t>(w.meta);
    BlockInfo *h_bl = (BlockInfo *)hm;
    uint8_t *h_codes = hm + o_codes;
    uint32_t *h_counts = (uint32_t *)(hm + bytes);
    const BlockInfo *d_bl = (const BlockInfo *)dm;
    const uint8_t *d_codes = dm + o_codes;
    const uint32_t nht = fill_tiles((uint32_t *)(hm + o_ht), B, [&](uint32_t b) { return (bl[b].n + HIST_TILE - 1) / HIST_TILE; });
    const uint32_t nst = fill_tiles((uint32_t *)(hm + o_st), B, [&](uint32_t b) { return (bl[b].n + stile - 1) / stile; });
    const uint32_t npt = fill_tiles((uint32_t *)(hm + o_pt), B, [&](uint32_t b) { return (bl[b].n + PART_TILE - 1) / PART_TILE; });
    memcpy(h_bl, bl.data(), B * sizeof(BlockInfo));
    HIPCHK(hipMemcpyAsync(dm, hm, o_codes, hipMemcpyHostToDevice, s));
    const Tiling thist{(const uint32_t *)(dm + o_ht), d_bl, B}, tsort{(const uint32_t *)(dm + o_st), d_bl, B};
    const Tiling tpart{(const uint32_t *)(dm + o_pt), d_bl, B};
    const uint16_t *blk = B > 1 ? P<uint16_t>(w.blk) : nullptr;

    HIPCHK(hipEventRecord(w.ev0, s));
    // histogram accumulators start at zero (every scan clears them; this covers a call that stopped half way)
    launch_fill_u8(P<uint8_t>(w.counts), 0, (size_t)B * MAXP * 256 * 4, s);
    if (B > 1) TIMED("fill_blk", N, launch_fill_blk(P<uint16_t>(w.blk), thist, nht, s));
    // 1. byte histograms -> dense symbol codes per block
    TIMED("byte_hist", N, launch_byte_hist(dT, thist, nht, P<uint32_t>(w.counts), s));
    TIMED("scan_digits", 256 * B,
          launch_scan_digits(P<uint32_t>(w.counts), B, 1, d_bl, 0, P<uint32_t>(w.tot), nullptr, s));
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(h_counts, w.tot.p, (size_t)B * 256 * 4, hipMemcpyDeviceToHost, s));
    SYNCCHK(s);
    // 56 key bits above the BWT byte: 8 symbols when a symbol needs 5-7 bits, 7 bytes of binary data, and as many
    // symbols as fit for small alphabets (14 of 4 bits, 18 of 3, 28 of 2, 56 of 1)
    uint32_t P0 = 1;
    for (uint32_t b = 0; b < B; b++) {
        const uint32_t *cnt = h_counts + (size_t)b * 256;
        uint8_t *code = h_codes + (size_t)b * 256;
        uint32_t sigma = 0;
        for (int c = 0; c < 256; c++) {
            code[c] = cnt[c] ? (uint8_t)sigma : 0;
            if (cnt[c]) sigma++;
        }
        const uint32_t wb = sigma <= 2 ? 1 : ceil_log2(sigma);
        bl[b].w = wb;
        bl[b].h0 = wb >= 5 ? std::min<uint32_t>(8, 56 / wb) : 56 / wb;
        P0 = std::max(P0, (bl[b].h0 * wb + 7) / 8);
        if (b == 0) {
            st.sigma = (int32_t)sigma;
            st.bits_per_symbol = (int32_t)wb;
            st.h0 = (int32_t)bl[b].h0;
        }
    }
    // digit histograms of the first keys: m-gram units (8-bit codes: one unit, from the byte counts) or, for 5-bit
    // codes, k_hist_T's tiles
    uint32_t maxn = 0;   // the largest block (the ISA scatter's runs follow its windows)
    for (uint32_t b = 0; b < B; b++) maxn = bl[b].n > maxn ? bl[b].n : maxn;
    bool plain_hist = false;
    const uint32_t ngu = fill_tiles((uint32_t *)(hm + o_gu), B, [&](uint32_t b) {
        if (!gram_block(bl[b].w)) {
            plain_hist = true;
            return 0u;
        }
        return bl[b].w == 8 ? 1u : (bl[b].n + GRAM_UNIT - 1) / GRAM_UNIT;
    });
    const Tiling tgram{(const uint32_t *)(dm + o_gu), d_bl, B};
    memcpy(h_bl, bl.data(), B * sizeof(BlockInfo));
    HIPCHK(hipMemcpyAsync(dm, hm, bytes, hipMemcpyHostToDevice, s));
    uint32_t *d_act = nullptr;
    if (feat) {   // router features: the byte histogram now, the active shares per round, the order-1 entropy last
        for (uint32_t b = 0; b < B; b++)
            for (int c = 0; c < 256; c++)
                feat[(size_t)b * ROUTER_NF + c] = std::sqrt((float)h_counts[(size_t)b * 256 + c] / (float)bl[b].n);
        grow(w.act, (size_t)ROUTER_ROUNDS * B * 4);
        d_act = P<uint32_t>(w.act);
        launch_fill_u8(P<uint8_t>(w.act), 0, (size_t)ROUTER_ROUNDS * B * 4, s);
    }

    // 2-4. The digit histograms of the first keys, the first sort, the groups after it and the ISA: a launch sequence
    // that depends only on the shapes and buffers below (the data travel in the metadata buffer), so that a call
    // repeating a recent call's shape replays it from an executable graph. Its lookbacks take their epochs from the
    // device epoch, which the sequence moves on first.
    uint64_t *K[2] = {P<uint64_t>(w.K0), P<uint64_t>(w.K1)};
    uint32_t *V[2] = {P<uint32_t>(w.V0), P<uint32_t>(w.V1)};
    uint32_t *SA = P<uint32_t>(w.SA), *ISA = P<uint32_t>(w.ISA);
    uint8_t *L = P<uint8_t>(w.L);
    uint32_t *depoch = w.counters + 15;
    if (++w.dev_epoch_calls >= EPOCH_HALF / EPOCH_STRIDE - 1) {   // the device epoch is about to come round again
        clear_status(w, s);
        w.dev_epoch_calls = 0;
    }
    auto first_sort = [&] {
        launch_epoch_bump(depoch, s);
        // 2. digit histograms of the first keys
        TIMED("hist_T", N, {
            if (plain_hist) launch_hist_T(dT, thist, nht, d_codes, P0, P<uint32_t>(w.counts), s);
            launch_hist_gram(dT, tgram, ngu, d_codes, P<uint32_t>(w.tot), P0, P<uint32_t>(w.counts), s);
        });
        TIMED("scan_digits", 256 * P0 * B,
              launch_scan_digits(P<uint32_t>(w.counts), B * P0, P0, d_bl, 1, nullptr, P<uint32_t>(w.dbase), s));
        LAUNCHCHK();
        // 3. the first sort: P0 stable passes per block over the digits above the BWT byte; the last writes the SA
        for (uint32_t p = 0; p < P0; p++) {
            SortArgs a{};
            a.cfg = scfg;
            a.from_t = p == 0;
            a.kin = p ? K[(p - 1) & 1] : nullptr;
            a.vin = p ? V[(p - 1) & 1] : nullptr;
            a.kout = K[p & 1];
            a.vout = p == P0 - 1 ? SA : V[p & 1];
            a.m = N;
            a.shift = 8 * p + 8;
            a.dbase = P<uint32_t>(w.dbase) + p * 256;
            a.dstride = P0 * 256;
            a.tiling = tsort;
            a.status = P<uint64_t>(w.status);
            a.counter = w.counters + 0;
            a.epoch = p;              // offsets 0..P0-1 from the device epoch; the partition below takes MAXP
            a.depoch = depoch;
            a.T = dT;
            a.codes = d_codes;
            TIMED(p == 0 ? "sort_pass_T" : "sort_pass", N, sort_pass_checked(a, nst, stile, s));
            LAUNCHCHK();
        }
        // 4. groups after the first sort; the BWT byte of every row is its key's low byte
        RegroupArgs g{};
        g.K = K[(P0 - 1) & 1];
        g.V = SA;
        g.pos = nullptr;
        g.m = N;
        g.kshift = 8;
        g.blk = blk;
        g.blocks = d_bl;
        g.SA = nullptr;
        g.ISA = ISA;
        g.pairs = K[P0 & 1];       // the other key buffer
        g.L = L;
        g.T = nullptr;
        g.carry = P<uint32_t>(w.carry);
        g.sh = P<uint32_t>(w.gh[0]);
        g.ss = P<uint32_t>(w.gs[0]);
        g.lh = P<uint32_t>(w.lh[0]);
        g.ls = P<uint32_t>(w.ls[0]);
        g.cnt = w.gcnt;
        launch_fill_u8(reinterpret_cast<uint8_t *>(w.gcnt), 0, 16, s);
        TIMED("regroup", N, launch_regroup(g, s));
        LAUNCHCHK();
        // ISA[suffix] = head: partition the pairs by suffix window inside each block (one keys-only radix pass),
        // then write window by window instead of scattering 4-byte writes over all of ISA
        TIMED("isa_part", N,
              launch_partition_pairs(K[P0 & 1], K[(P0 - 1) & 1], N, tpart, npt, P<uint64_t>(w.status),
                                     w.counters + 4, MAXP, depoch, s));
        LAUNCHCHK();
        TIMED("isa_scatter", N, launch_isa_scatter(K[(P0 - 1) & 1], N, maxn, blk, d_bl, ISA, s));
        LAUNCHCHK();
    };
    if (feat) {   // the router's path: its per-round counts are set up around this sequence, so no graph
        first_sort();
    } else {
        const std::vector<uint64_t> key = {
            B, N, P0, (uint64_t)scfg, nht, nst, npt, ngu, (uint64_t)plain_hist, maxn, (uintptr_t)w.tot.p, (uintptr_t)dT,
            (uintptr_t)dm, (uintptr_t)w.counts.p,
            (uintptr_t)w.dbase.p, (uintptr_t)w.K0.p, (uintptr_t)w.K1.p, (uintptr_t)w.V0.p, (uintptr_t)w.V1.p,
            (uintptr_t)w.SA.p, (uintptr_t)w.ISA.p, (uintptr_t)w.L.p, (uintptr_t)w.carry.p, (uintptr_t)w.gh[0].p,
            (uintptr_t)w.gs[0].p, (uintptr_t)w.lh[0].p, (uintptr_t)w.ls[0].p, (uintptr_t)w.status.p,
            (uintptr_t)w.counters, (uintptr_t)w.gcnt, (uintptr_t)blk};
        run_graph(w.fwd_graph, key, s,
                  {"hist_T", "scan_digits", "sort_pass_T", "sort_pass", "regroup", "isa_part", "isa_scatter"},
                  first_sort);
    }
    st.passes[0] = (int32_t)P0;
    st.active[0] = N;

    // 5. doubling rounds over the groups of two or more (SA ranges). Small groups sort inside one workgroup each
    // (k_seg_round); groups of more than SEG_GROUP rows take the global path (round keys, radix passes, regroup).
    grow(w.go, GMAX * 4);
    grow(w.tstart, (NN / SEG_TILE + 2) * 4);
    grow(w.eh, (NN / SEG_TILE + 2) * SEG_TILE * 4);
    grow(w.es, (NN / SEG_TILE + 2) * SEG_TILE * 4);
    grow(w.ecnt, (NN / SEG_TILE + 2) * 4);
    grow(w.eoff, (NN / SEG_TILE + 2) * 4);
    grow(w.tmax, (NN / SEG_TILE + 2) * 4);
    grow(w.sec, NN * 4);
    grow(w.sufl, NN * 4);
    grow(w.lrow, NN);
    grow(w.lo, GMAX * 8);
    uint32_t *gcnt = w.gcnt;          // [small groups, large groups] of the list being built
    const uint32_t rank_max = tune.seg_rank;
    std::vector<uint32_t> h_lh, h_ls, h_lo;
    // The list counters, and the large list itself. While large groups remain, the first LIST_PEEK heads and sizes
    // come back with the counters (one wait instead of two); a longer list takes a second copy.
    bool peek = true;
    auto read_lists = [&](int c, uint32_t &Gs, uint32_t &Gl) {
        const size_t np = peek ? std::min<size_t>(LIST_PEEK, std::min(w.lh[c].cap, w.ls[c].cap) / 4) : 0;
        HIPCHK(hipMemcpyAsync(w.h_gcnt, gcnt, 16, hipMemcpyDeviceToHost, s));
        if (np) {
            HIPCHK(hipMemcpyAsync(w.h_peek, w.lh[c].p, np * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(w.h_peek + LIST_PEEK, w.ls[c].p, np * 4, hipMemcpyDeviceToHost, s));
        }
        SYNCCHK(s);
        Gs = w.h_gcnt[0];
        Gl = w.h_gcnt[1];
        h_lh.resize(Gl);
        h_ls.resize(Gl);
        if (Gl && Gl <= np) {
            memcpy(h_lh.data(), w.h_peek, (size_t)Gl * 4);
            memcpy(h_ls.data(), w.h_peek + LIST_PEEK, (size_t)Gl * 4);
        } else if (Gl) {
            HIPCHK(hipMemcpyAsync(h_lh.data(), w.lh[c].p, (size_t)Gl * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(h_ls.data(), w.ls[c].p, (size_t)Gl * 4, hipMemcpyDeviceToHost, s));
            SYNCCHK(s);
        }
        peek = Gl > 0;
    };
    auto zero_counts = [&] { launch_fill_u8(reinterpret_cast<uint8_t *>(gcnt), 0, 16, s); };
    int cur = 0, round = 0;
    uint32_t Gs = 0, Gl = 0;
    read_lists(0, Gs, Gl);   // the lists the first regroup wrote
    const Tiling flat{nullptr, d_bl, 1};
    while (Gs + Gl > 0) {
        round++;
        if (round >= KBWT_MAX_ROUNDS) fail(hipErrorUnknown, "doubling did not converge");
        if (d_act && round <= ROUTER_ROUNDS) {   // rows of each block still in groups at the start of this round
            uint32_t *a = d_act + (size_t)(round - 1) * B;
            launch_count_active(P<uint32_t>(w.gh[cur]), P<uint32_t>(w.gs[cur]), Gs, blk, a, s);
            launch_count_active(P<uint32_t>(w.lh[cur]), P<uint32_t>(w.ls[cur]), Gl, blk, a, s);
            LAUNCHCHK();
        }
        uint32_t Es = 0, ml = 0;
        // every key of the round is gathered before any rank changes: the small groups' offsets and the large groups'
        // keys first (neither writes a rank), then one wait for the small groups' rows and the digits that vary
        if (Gs) {
            TIMED("scan_groups", Gs,
                  launch_scan_groups(P<uint32_t>(w.gs[cur]), Gs, P<uint32_t>(w.go), w.m_out, P<uint32_t>(w.tstart),
                                     P<uint64_t>(w.st_sum), w.counters + 3, next_epoch(w, s), s));
            LAUNCHCHK();
            HIPCHK(hipMemcpyAsync(w.h_m, w.m_out, 4, hipMemcpyDeviceToHost, s));
        }
        Shifts sh{};
        if (Gl) {
            // the large groups in increasing head order (regroup's head scan runs over SA positions), then their rows
            std::vector<uint32_t> ord(Gl);
            for (uint32_t g = 0; g < Gl; g++) ord[g] = g;
            std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return h_lh[x] < h_lh[y]; });
            h_lo.resize(2 * (size_t)Gl);
            for (uint32_t k = 0; k < Gl; k++) {
                h_lo[k] = h_lh[ord[k]];
                h_lo[Gl + k] = ml;
                ml += h_ls[ord[k]];
            }
            HIPCHK(hipMemcpyAsync(w.lo.p, h_lo.data(), (size_t)2 * Gl * 4, hipMemcpyHostToDevice, s));
            TIMED("large_elems", ml,
                  launch_large_elems(P<uint32_t>(w.lo), P<uint32_t>(w.lo) + Gl, Gl, ml, P<uint32_t>(w.pos[0]),
                                     P<uint32_t>(w.grp[0]), s));
            HIPCHK(hipMemcpyAsync(w.orand, w.h_orand_init, 16, hipMemcpyHostToDevice, s));
            TIMED("round_keys", ml,
                  launch_round_keys(P<uint32_t>(w.pos[0]), P<uint32_t>(w.grp[0]), ml, SA, ISA, N, blk, d_bl, (uint32_t)round,
                                    K[0], V[0], w.orand, s));
            LAUNCHCHK();
            HIPCHK(hipMemcpyAsync(w.h_orand, w.orand, 16, hipMemcpyDeviceToHost, s));
        }
        SYNCCHK(s);
        if (Gs) Es = *w.h_m;
        if (Gl) {
            const uint64_t vary = w.h_orand[0] ^ w.h_orand[1];
            for (uint32_t d = 0; d < 8; d++)
                if ((vary >> (8 * d)) & 0xFFull) sh.s[sh.P++] = 8 * d;
        }
        SegRoundArgs sa{};
        sa.gh = P<uint32_t>(w.gh[cur]);
        sa.go = P<uint32_t>(w.go);
        sa.tstart = P<uint32_t>(w.tstart);
        sa.G = Gs;
        sa.E = Es;
        sa.SA = SA;
        sa.ISA = ISA;
        sa.sec = P<uint32_t>(w.sec);
        sa.sufl = P<uint32_t>(w.sufl);
        sa.L = L;
        sa.lrow = P<uint8_t>(w.lrow);
        sa.T = dT;
        sa.blk = blk;
        sa.blocks = d_bl;
        sa.n = N;
        sa.round = (uint32_t)round;
        sa.nh = P<uint32_t>(w.eh);
        sa.ns = P<uint32_t>(w.es);
        sa.ncnt = P<uint32_t>(w.ecnt);
        sa.tmax = P<uint32_t>(w.tmax);
        sa.rank_max = rank_max;
        if (Gs) TIMED("seg_gather", Es, launch_seg_round(0, sa, s));
        // sort and split; the next round's lists are built from zero
        zero_counts();
        if (Gs) {
            // the split: direct ranking for tiles whose splitting groups are small, the LDS radix sort for the rest
            TIMED("seg_sort", Es, launch_seg_round(2, sa, s); launch_seg_round(1, sa, s));
            // the tiles' emissions into the next small list; its length goes to gcnt[0], where the large groups'
            // subgroups are appended after it
            const uint32_t nt = (Es + SEG_TILE - 1) / SEG_TILE;
            TIMED("emit_compact", Es,
                  launch_scan_groups(P<uint32_t>(w.ecnt), nt, P<uint32_t>(w.eoff), gcnt, nullptr,
                                     P<uint64_t>(w.st_sum), w.counters + 3, next_epoch(w, s), s);
                  launch_compact_emits(P<uint32_t>(w.eh), P<uint32_t>(w.es), P<uint32_t>(w.ecnt), P<uint32_t>(w.eoff),
                                       nt, P<uint32_t>(w.gh[cur ^ 1]), P<uint32_t>(w.gs[cur ^ 1]), s));
        }
        LAUNCHCHK();
        if (Gl) {
            int src = 0;
            if (sh.P) {
                const uint32_t g2 = std::max(1u, std::min<uint32_t>(HIST_GRID, (ml + 8191) / 8192));
                TIMED("hist_keys", ml, launch_hist_keys(K[0], ml, sh, P<uint32_t>(w.counts), g2, s));
                TIMED("scan_digits", 256 * sh.P,
                      launch_scan_digits(P<uint32_t>(w.counts), sh.P, MAXP, d_bl, 0, nullptr, P<uint32_t>(w.dbase), s));
                LAUNCHCHK();
                const uint32_t ntl = (ml + stile - 1) / stile;
                for (uint32_t p = 0; p < sh.P; p++) {
                    SortArgs a{};
                    a.cfg = scfg;
                    a.from_t = false;
                    a.kin = K[src];
                    a.vin = V[src];
                    a.kout = K[src ^ 1];
                    a.vout = V[src ^ 1];
                    a.m = ml;
                    a.shift = sh.s[p];
                    a.dbase = P<uint32_t>(w.dbase) + p * 256;
                    a.dstride = 0;
                    a.tiling = flat;
                    a.status = P<uint64_t>(w.status);
                    a.counter = w.counters + 0;
                    a.epoch = next_epoch(w, s);
                    TIMED("sort_pass", ml, sort_pass_checked(a, ntl, stile, s));
                    LAUNCHCHK();
                    src ^= 1;
                }
            }
            RegroupArgs g{};
            g.K = K[src];
            g.V = V[src];
            g.pos = P<uint32_t>(w.pos[0]);
            g.m = ml;
            g.kshift = 0;
            g.blk = nullptr;
            g.SA = SA;
            g.ISA = ISA;
            g.L = L;
            g.T = dT;
            g.blocks = d_bl;
            g.carry = P<uint32_t>(w.carry);
            g.sh = P<uint32_t>(w.gh[cur ^ 1]);
            g.ss = P<uint32_t>(w.gs[cur ^ 1]);
            g.lh = P<uint32_t>(w.lh[cur ^ 1]);
            g.ls = P<uint32_t>(w.ls[cur ^ 1]);
            g.cnt = gcnt;
            TIMED("regroup", ml, launch_regroup(g, s));
            LAUNCHCHK();
        }
        st.passes[round] = (int32_t)sh.P;
        st.active[round] = (int64_t)Es + ml;
        cur ^= 1;
        read_lists(cur, Gs, Gl);
    }
    st.rounds = round;

    // 6. BWT bytes (the row bytes shifted around each block's primary row) and sampled indexes
    if (feat) router_features_tail(w, dT, d_bl, bl, d_act, feat, s);
    TIMED("bwt", N, launch_bwt(dT, L, ISA, blk, d_bl, dU, N, s));
    TIMED("samples", nsamp, launch_samples(ISA, d_bl, B, nsamp, P<int32_t>(w.samples), s));
    LAUNCHCHK();
    HIPCHK(hipEventRecord(w.ev1, s));
    HIPCHK(hipMemcpyAsync(I, w.samples.p, (size_t)nsamp * 4, hipMemcpyDeviceToHost, s));
    SYNCCHK(s);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, w.ev0, w.ev1));
    st.ms_fwd = ms;
    prof_collect(w);
}

// ------------------------------------------------------------------ inverse

// The inverse's shapes for a batch: bl gets roff, spoff and nreg; one splitter row in D = 2^dshift; the first cap
// bytes of each sublist are staged during the counting walk (the walk's time is its longest sublist, about
// D * ln(splitters) steps, times the load latency), in a chunk-major stage of cap / 16 rows of 16-byte chunks, one per
// splitter (sstride >= S of them), at most stage_max bytes in all.
struct InvPlan {
    uint32_t B, N, S, dshift, cap, iters, nht, nlt, nwt;
    size_t sstride;
    size_t o_ht, o_lt, o_wt, meta_bytes;   // metadata layout: blocks | hist tiles | lf tiles | walk tiles
};
InvPlan plan_inverse(BlockInfo *bl, uint32_t B, uint32_t N, uint32_t dshift, uint32_t cap, size_t stage_max) {
    InvPlan p{};
    p.B = B;
    p.N = N;
    p.dshift = dshift;
    const uint32_t dmask = (1u << dshift) - 1u;
    for (uint32_t b = 0; b < B; b++) {
        bl[b].roff = bl[b].off + b;
        bl[b].spoff = p.S;
        bl[b].nreg = (bl[b].n >> dshift) + 1;
        p.S += bl[b].nreg + ((bl[b].P & dmask) != 0 ? 1u : 0u);
    }
    p.sstride = ((size_t)p.S + 63) & ~(size_t)63;
    if ((size_t)cap * p.sstride > stage_max) cap = (uint32_t)std::max<size_t>(16, stage_max / p.sstride);
    p.cap = std::max(16u, cap & ~15u);           // the counting walk stages whole 16-step chunks
    p.iters = rank_jump_iters(p.S);
    p.o_ht = align16(B * sizeof(BlockInfo));
    p.o_lt = p.o_ht + (B + 1) * 4;
    p.o_wt = p.o_lt + (B + 1) * 4;
    p.meta_bytes = p.o_wt + (B + 1) * 4;
    return p;
}
// The metadata of the plan into host memory hm (meta_bytes); also sets the tile counts.
void inverse_meta(InvPlan &p, const BlockInfo *bl, uint8_t *hm) {
    const uint32_t B = p.B, dmask = (1u << p.dshift) - 1u;
    memcpy(hm, bl, B * sizeof(BlockInfo));
    p.nht = fill_tiles((uint32_t *)(hm + p.o_ht), B, [&](uint32_t b) { return (bl[b].n + HIST_TILE - 1) / HIST_TILE; });
    p.nlt = fill_tiles((uint32_t *)(hm + p.o_lt), B, [&](uint32_t b) { return (bl[b].n + LF_TILE - 1) / LF_TILE; });
    p.nwt = fill_tiles((uint32_t *)(hm + p.o_wt), B, [&](uint32_t b) {
        return (bl[b].nreg + ((bl[b].P & dmask) != 0 ? 1u : 0u) + WALK_TILE - 1) / WALK_TILE;
    });
}

// Device buffers of one inverse call. Lookback epochs: depoch (a device word a bump kernel moves on, for launch
// sequences a graph replays) or, when depoch is null, the host-side epoch.
struct InvBufs {
    uint32_t *lf;            // N + 64 words
    uint64_t *pair;          // N + 64
    uint64_t *link[2];       // S + 1 each
    uint32_t *len, *cont;    // S + 1 each
    uint8_t *stage;          // cap * sstride
    uint64_t *status;        // (N / LF_TILE + B + 1) * 256
    uint32_t *counts;        // B * MAXP * 256 (cleared by the sequence)
    uint32_t *cbase;         // B * 256
    uint32_t *bad;           // B
    uint16_t *blk;           // N (B > 1 only)
    uint8_t *meta;           // meta_bytes
    uint32_t *tile_counter;  // zero between launches
    uint32_t *depoch;
    uint32_t epoch;
};
size_t inv_status_words(uint32_t N, uint32_t B) { return ((size_t)std::max<uint32_t>(N, 64) + LF_TILE - 1) / LF_TILE + B; }

// The launch sequence of one inverse (every kernel on stream s; wp: the workspace whose events time the launches
// while profiling, or null).
void inverse_launches(const InvPlan &p, const InvBufs &bf, const uint8_t *dL, uint8_t *dOut, Workspace *wp,
                      hipStream_t s) {
    const uint32_t B = p.B, N = p.N, S = p.S;
    auto tm = [&](const char *name, uint64_t elems, auto fn) {
        if (wp) timed(*wp, s, name, elems, fn);
        else fn();
    };
    const BlockInfo *d_bl = (const BlockInfo *)bf.meta;
    const Tiling thist{(const uint32_t *)(bf.meta + p.o_ht), d_bl, B};
    const Tiling tlf{(const uint32_t *)(bf.meta + p.o_lt), d_bl, B};
    const Tiling twalk{(const uint32_t *)(bf.meta + p.o_wt), d_bl, B};
    if (bf.depoch) launch_epoch_bump(bf.depoch, s);
    launch_fill_u8(reinterpret_cast<uint8_t *>(bf.counts), 0, (size_t)B * MAXP * 256 * 4, s);
    if (B > 1) tm("fill_blk", N, [&] { launch_fill_blk(bf.blk, thist, p.nht, s); });
    tm("byte_hist", N, [&] { launch_byte_hist(dL, thist, p.nht, bf.counts, s); });
    tm("scan_digits", 256 * B, [&] { launch_scan_digits(bf.counts, B, 1, d_bl, 2, nullptr, bf.cbase, s); });
    tm("lf_pass", N, [&] {
        launch_lf_pass(dL, tlf, p.nlt, bf.cbase, bf.lf, bf.status, bf.tile_counter, bf.depoch ? 0u : bf.epoch,
                       bf.depoch, s);
    });
    tm("lf_pairs", N, [&] { launch_lf_pairs(bf.lf, dL, tlf, p.nlt, p.dshift, bf.pair, s); });
    tm("walk_count", N, [&] {
        launch_walk_count(bf.pair, twalk, p.nwt, p.dshift, S, bf.cbase, bf.link[0], bf.len, bf.stage, p.sstride, p.cap,
                          bf.cont, s);
    });
    int cur = 0;
    for (uint32_t it = 0; it < p.iters; it++) {
        tm("rank_jump", S, [&] { launch_rank_jump(bf.link[cur], bf.link[cur ^ 1], S, s); });
        cur ^= 1;
    }
    launch_check_lists(bf.link[cur], d_bl, B, S, bf.bad, s);
    tm("walk_emit", N, [&] {
        launch_walk_emit(bf.lf, twalk, p.nwt, p.dshift, S, bf.link[cur], bf.len, bf.cbase, bf.stage, p.sstride, p.cap,
                         bf.cont, bf.bad, dOut, s);
    });
}

// bl: one BlockInfo per block with off, n (>= 2) and P (primary index) filled in; packed BWTs dL, output dOut (may
// be dL).
void inverse(Workspace &w, const uint8_t *dL, uint8_t *dOut, std::vector<BlockInfo> &bl, uint32_t N, hipStream_t s) {
    const uint32_t B = (uint32_t)bl.size();
    kbwt_stats &st = g_stats;
    memset(&st, 0, sizeof st);   // the stats describe this call alone: the forward's fields stay zero
    st.n = (int32_t)N;
    st.blocks = (int32_t)B;
    const uint32_t dshift = tuning_for(N).inv_dshift;
    InvPlan p = plan_inverse(bl.data(), B, N, dshift, env_u32("KBWT_INV_CAP", 16u << dshift, 1, 1u << 16), STAGE_MAX);
    st.splitters = (int32_t)p.S;
    const size_t NN = std::max<size_t>(N, 64), S = p.S;
    grow(w.V0, NN * 4);                        // lf
    grow(w.pos[0], (S + 1) * 4);               // sublist lengths
    grow(w.pos[1], (S + 1) * 4);               // continuation rows
    grow(w.K0, (S + 1) * 8);                   // links (ping)
    grow(w.K1, (S + 1) * 8);                   // links (pong)
    grow(w.pair, NN * 8);                      // two LF steps per row
    grow(w.stage, (size_t)p.cap * p.sstride);
    grow(w.status, inv_status_words(N, B) * 256 * 8, true);
    grow(w.counts, (size_t)B * MAXP * 256 * 4, true);
    grow(w.cbase, (size_t)B * 256 * 4);
    grow(w.bad, (size_t)B * 4);
    if (B > 1) grow(w.blk, NN * 2);
    uint8_t *hm = (uint8_t *)host_staging(w, align16(p.meta_bytes));
    grow(w.meta, p.meta_bytes);
    inverse_meta(p, bl.data(), hm);
    HIPCHK(hipMemcpyAsync(w.meta.p, hm, p.meta_bytes, hipMemcpyHostToDevice, s));
    InvBufs bf{};
    bf.lf = P<uint32_t>(w.V0);
    bf.pair = P<uint64_t>(w.pair);
    bf.link[0] = P<uint64_t>(w.K0);
    bf.link[1] = P<uint64_t>(w.K1);
    bf.len = P<uint32_t>(w.pos[0]);
    bf.cont = P<uint32_t>(w.pos[1]);
    bf.stage = P<uint8_t>(w.stage);
    bf.status = P<uint64_t>(w.status);
    bf.counts = P<uint32_t>(w.counts);
    bf.cbase = P<uint32_t>(w.cbase);
    bf.bad = P<uint32_t>(w.bad);
    bf.blk = B > 1 ? P<uint16_t>(w.blk) : nullptr;
    bf.meta = P<uint8_t>(w.meta);
    bf.tile_counter = w.counters + 2;
    bf.depoch = w.counters + 15;

    HIPCHK(hipEventRecord(w.ev0, s));
    // the launch sequence depends only on these shapes and buffers (the data travel in the metadata buffer)
    const std::vector<uint64_t> key = {
        B, N, S, p.dshift, p.cap, p.sstride, p.nht, p.nlt, p.nwt, (uintptr_t)dL, (uintptr_t)dOut, (uintptr_t)w.meta.p,
        (uintptr_t)w.V0.p, (uintptr_t)w.K0.p, (uintptr_t)w.K1.p, (uintptr_t)w.pos[0].p, (uintptr_t)w.pos[1].p,
        (uintptr_t)w.stage.p, (uintptr_t)w.status.p, (uintptr_t)w.counts.p, (uintptr_t)w.cbase.p,
        (uintptr_t)w.bad.p, (uintptr_t)w.blk.p, (uintptr_t)w.counters, (uintptr_t)w.pair.p};
    // every call moves the device epoch on by EPOCH_STRIDE inside [EPOCH_HALF, 2 * EPOCH_HALF); before it wraps, the
    // status words are cleared (outside any graph)
    if (++w.dev_epoch_calls >= EPOCH_HALF / EPOCH_STRIDE - 1) {
        clear_status(w, s);
        w.dev_epoch_calls = 0;
    }
    run_graph(w.inv_graph, key, s,
              {"fill_blk", "byte_hist", "scan_digits", "lf_pass", "lf_pairs", "walk_count", "rank_jump", "walk_emit"},
              [&] { inverse_launches(p, bf, dL, dOut, &w, s); });
    LAUNCHCHK();   // a block whose string is no BWT decodes to zeros inside the emission: nothing to read back
    HIPCHK(hipEventRecord(w.ev1, s));
    SYNCCHK(s);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, w.ev0, w.ev1));
    st.ms_inv = ms;
    prof_collect(w);
}

// ---- the inverse that allocates nothing: kbwt_dev_unbwt_aux_noalloc
//
// Every device buffer of one call comes from `scratch`, a file-scope device array of NOALLOC_BYTES (about 100 MiB),
// sized for blocks of up to NOALLOC_MAX_N bytes with splitters at least every 8 rows and at most 8 bytes of staged
// sublist heads per input byte; each device that runs the library holds it until the process ends. In host code the
// array's name is only a stand-in: its device address comes from hipGetSymbolAddress once per device (with that
// device current) and only that address is used. Metadata goes up from the caller's stack; the launch sequence is the
// inverse's own (inverse_launches) with lookback epochs counted on the host.
// Calls that overlap on one device share the array and race.
constexpr uint32_t NOALLOC_MAX_N = 1u << 22;
constexpr uint32_t NOALLOC_MIN_DSHIFT = 3;
constexpr size_t NOALLOC_SMAX = NOALLOC_MAX_N / 8 + 2;
constexpr size_t NOALLOC_STAGE = (size_t)8 * NOALLOC_MAX_N;
constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_LF = 0;
constexpr size_t NA_PAIR = NA_LF + al256(((size_t)NOALLOC_MAX_N + 64) * 4);
constexpr size_t NA_LINK0 = NA_PAIR + al256(((size_t)NOALLOC_MAX_N + 64) * 8);
constexpr size_t NA_LINK1 = NA_LINK0 + al256((NOALLOC_SMAX + 1) * 8);
constexpr size_t NA_LEN = NA_LINK1 + al256((NOALLOC_SMAX + 1) * 8);
constexpr size_t NA_CONT = NA_LEN + al256((NOALLOC_SMAX + 1) * 4);
constexpr size_t NA_STAGE = NA_CONT + al256((NOALLOC_SMAX + 1) * 4);
constexpr size_t NA_STATUS = NA_STAGE + al256(NOALLOC_STAGE);
constexpr size_t NA_STATUS_BYTES = al256((((size_t)NOALLOC_MAX_N + LF_TILE - 1) / LF_TILE + 2) * 256 * 8);
constexpr size_t NA_COUNTS = NA_STATUS + NA_STATUS_BYTES;
constexpr size_t NA_CBASE = NA_COUNTS + al256((size_t)MAXP * 256 * 4);
constexpr size_t NA_BAD = NA_CBASE + al256(256 * 4);
constexpr size_t NA_META = NA_BAD + 256;
constexpr size_t NA_COUNTERS = NA_META + 256;
constexpr size_t NA_T = NA_COUNTERS + 256;     // an input or output that is not 16-byte aligned goes through here
constexpr size_t NOALLOC_BYTES = NA_T + al256((size_t)NOALLOC_MAX_N + 64);

static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

struct NoAllocDevice {
    std::mutex mu;
    uint8_t *base = nullptr;   // the device address of scratch on this device
    uint32_t epoch = 1;        // host-side lookback epochs of the calls on this device
};
NoAllocDevice g_noalloc[64];

// The device address of scratch on the current device (found on the first call there, whose stream also clears the
// lookback status words), and this call's lookback epoch (the status words are cleared again before epochs wrap).
uint8_t *noalloc_scratch(hipStream_t s, uint32_t &epoch) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) fail(hipErrorInvalidDevice, "hipGetDevice");
    NoAllocDevice &d = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(d.mu);
    bool clear = false;
    if (!d.base) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        d.base = static_cast<uint8_t *>(p);
        clear = true;
    }
    if (d.epoch >= EPOCH_HALF - 1) {
        d.epoch = 1;
        clear = true;
    }
    if (clear) {
        launch_fill_u8(d.base + NA_STATUS, 0, NA_STATUS_BYTES, s);
        LAUNCHCHK();
    }
    epoch = d.epoch++;
    return d.base;
}

void inverse_noalloc(const uint8_t *dT, uint8_t *dU, uint32_t n, uint32_t primary, hipStream_t s) {
    kbwt_stats &st = g_stats;
    memset(&st, 0, sizeof st);
    st.n = (int32_t)n;
    st.blocks = 1;
    if (n <= 1) {
        if (n == 1 && dU != dT) HIPCHK(hipMemcpyAsync(dU, dT, 1, hipMemcpyDeviceToDevice, s));
        SYNCCHK(s);
        return;
    }
    uint32_t epoch = 0;
    uint8_t *base = noalloc_scratch(s, epoch);
    BlockInfo bl{};
    bl.off = 0;
    bl.n = n;
    bl.P = primary;
    const uint32_t dshift = std::max(NOALLOC_MIN_DSHIFT, tuning_for(n).inv_dshift);
    InvPlan p = plan_inverse(&bl, 1, n, dshift, 16u << dshift, NOALLOC_STAGE);
    st.splitters = (int32_t)p.S;
    alignas(16) uint8_t hm[256];
    static_assert(sizeof(BlockInfo) <= 64, "metadata of one block fits the stack buffer");
    inverse_meta(p, &bl, hm);
    InvBufs bf{};
    bf.lf = reinterpret_cast<uint32_t *>(base + NA_LF);
    bf.pair = reinterpret_cast<uint64_t *>(base + NA_PAIR);
    bf.link[0] = reinterpret_cast<uint64_t *>(base + NA_LINK0);
    bf.link[1] = reinterpret_cast<uint64_t *>(base + NA_LINK1);
    bf.len = reinterpret_cast<uint32_t *>(base + NA_LEN);
    bf.cont = reinterpret_cast<uint32_t *>(base + NA_CONT);
    bf.stage = base + NA_STAGE;
    bf.status = reinterpret_cast<uint64_t *>(base + NA_STATUS);
    bf.counts = reinterpret_cast<uint32_t *>(base + NA_COUNTS);
    bf.cbase = reinterpret_cast<uint32_t *>(base + NA_CBASE);
    bf.bad = reinterpret_cast<uint32_t *>(base + NA_BAD);
    bf.blk = nullptr;
    bf.meta = base + NA_META;
    bf.tile_counter = reinterpret_cast<uint32_t *>(base + NA_COUNTERS);
    bf.depoch = nullptr;
    bf.epoch = epoch;
    // the walks read only the LF table, so the output may overwrite the input; the kernels take 16-byte aligned rows
    const bool direct = ((uintptr_t)dT & 15) == 0 && (dT == dU || ((uintptr_t)dU & 15) == 0);
    const uint8_t *src = direct ? dT : base + NA_T;
    uint8_t *dst = direct ? dU : base + NA_T;
    if (!direct) HIPCHK(hipMemcpyAsync(base + NA_T, dT, n, hipMemcpyDeviceToDevice, s));
    launch_fill_u8(base + NA_COUNTERS, 0, 64, s);   // the tile counter starts at zero whatever a failed call left
    HIPCHK(hipMemcpyAsync(bf.meta, hm, p.meta_bytes, hipMemcpyHostToDevice, s));
    inverse_launches(p, bf, src, dst, nullptr, s);   // a string that is no BWT decodes to zeros
    LAUNCHCHK();
    if (!direct) HIPCHK(hipMemcpyAsync(dU, base + NA_T, n, hipMemcpyDeviceToDevice, s));
    SYNCCHK(s);
}

// ------------------------------------------------------------------ batch plumbing

bool forward_args_ok(int32_t n, int32_t r) { return n >= 0 && r >= 2 && (r & (r - 1)) == 0; }
bool inverse_args_ok(int32_t n, int32_t r) { return n >= 0 && (r == n || (r >= 2 && (r & (r - 1)) == 0)); }
inline uint32_t nsamples(int32_t n, int32_t r) { return n > 0 ? (uint32_t)((n - 1) / r + 1) : 1u; }

// Validates a batch: increasing, non-overlapping blocks, at most 65535, total length < 2^31.
bool batch_ok(const kbwt_seg *segs, int32_t count, bool fwd) {
    if (!segs || count < 0 || count > 65535) return false;
    int64_t end = 0, total = 0;
    for (int32_t b = 0; b < count; b++) {
        const kbwt_seg &g = segs[b];
        if (g.offset < end) return false;
        if (fwd ? !forward_args_ok(g.n, g.r) : !inverse_args_ok(g.n, g.r)) return false;
        end = g.offset + g.n;
        total += g.n;
    }
    return total < ((int64_t)1 << 31) - 64;
}

// True when every block has n >= 2, they sit back to back, and the first starts 16-byte aligned.
bool packed(const uint8_t *base, const kbwt_seg *segs, int32_t count) {
    int64_t at = -1;
    for (int32_t b = 0; b < count; b++) {
        if (segs[b].n <= 1) return false;
        if (at >= 0 && segs[b].offset != at) return false;
        at = segs[b].offset + segs[b].n;
    }
    return count == 0 || (((uintptr_t)(base + segs[0].offset)) & 15) == 0;
}

// Forward on a batch whose bytes sit in device memory (dT, dU device pointers; I host).
void forward_batch_dev(Workspace &w, const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count, int32_t *I,
                       hipStream_t s, float *feat = nullptr) {
    std::vector<uint32_t> soff(count + 1, 0);
    for (int32_t b = 0; b < count; b++) soff[b + 1] = soff[b] + nsamples(segs[b].n, segs[b].r);
    std::vector<BlockInfo> bl;
    std::vector<int32_t> which;
    uint32_t N = 0, nsamp = 0;
    for (int32_t b = 0; b < count; b++) {
        if (segs[b].n <= 1) continue;
        BlockInfo bi{};
        bi.off = N;
        bi.n = (uint32_t)segs[b].n;
        bi.r = (uint32_t)segs[b].r;
        bi.soff = nsamp;
        bl.push_back(bi);
        which.push_back(b);
        N += bi.n;
        nsamp += nsamples(segs[b].n, segs[b].r);
    }
    // the BWT reads T at random while writing U: in place needs a copy, and so does a batch that is not packed
    const bool direct = packed(dT, segs, count) && dT != dU;
    std::vector<int32_t> Id(std::max<uint32_t>(nsamp, 1));
    if (!bl.empty()) {
        const uint8_t *src = dT + segs[0].offset;
        uint8_t *dst = dU + segs[0].offset;
        if (!direct) {
            grow(w.T, (size_t)N + 64);
            grow(w.U, (size_t)N + 64);
            for (size_t k = 0; k < bl.size(); k++)
                HIPCHK(hipMemcpyAsync(P<uint8_t>(w.T) + bl[k].off, dT + segs[which[k]].offset, bl[k].n,
                                      hipMemcpyDeviceToDevice, s));
            src = P<uint8_t>(w.T);
            dst = P<uint8_t>(w.U);
        }
        std::vector<float> fi(feat ? bl.size() * ROUTER_NF : 0);
        forward(w, src, dst, bl, N, Id.data(), nsamp, s, feat ? fi.data() : nullptr);
        if (feat)
            for (size_t k = 0; k < bl.size(); k++)
                memcpy(feat + (size_t)which[k] * ROUTER_NF, fi.data() + k * ROUTER_NF, ROUTER_NF * 4);
        if (!direct)
            for (size_t k = 0; k < bl.size(); k++)
                HIPCHK(hipMemcpyAsync(dU + segs[which[k]].offset, P<uint8_t>(w.U) + bl[k].off, bl[k].n,
                                      hipMemcpyDeviceToDevice, s));
    }
    for (int32_t b = 0; b < count; b++)
        if (segs[b].n <= 1) {
            if (segs[b].n == 1)
                HIPCHK(hipMemcpyAsync(dU + segs[b].offset, dT + segs[b].offset, 1, hipMemcpyDeviceToDevice, s));
            I[soff[b]] = segs[b].n;
            if (feat) {   // nothing to sort: the histogram of the one byte, if any
                float *f = feat + (size_t)b * ROUTER_NF;
                memset(f, 0, ROUTER_NF * 4);
                if (segs[b].n == 1) {
                    uint8_t c = 0;
                    HIPCHK(hipMemcpyAsync(&c, dT + segs[b].offset, 1, hipMemcpyDeviceToHost, s));
                    SYNCCHK(s);
                    f[c] = 1.f;
                }
            }
        }
    for (size_t k = 0; k < bl.size(); k++)
        memcpy(I + soff[which[k]], Id.data() + bl[k].soff, nsamples((int32_t)bl[k].n, (int32_t)bl[k].r) * 4);
    SYNCCHK(s);
}

// Inverse on a batch in device memory; dT == dU allowed.
void inverse_batch_dev(Workspace &w, const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                       const int32_t *I, hipStream_t s) {
    std::vector<uint32_t> soff(count + 1, 0);
    for (int32_t b = 0; b < count; b++) soff[b + 1] = soff[b] + nsamples(segs[b].n, segs[b].r);
    std::vector<BlockInfo> bl;
    std::vector<int32_t> which;
    uint32_t N = 0;
    for (int32_t b = 0; b < count; b++) {
        if (segs[b].n <= 1) continue;
        BlockInfo bi{};
        bi.off = N;
        bi.n = (uint32_t)segs[b].n;
        bi.P = (uint32_t)I[soff[b]];
        bl.push_back(bi);
        which.push_back(b);
        N += bi.n;
    }
    // the walks read only the LF table, so the output may overwrite the input
    const bool direct = packed(dT, segs, count) && (dT == dU || packed(dU, segs, count));
    if (!bl.empty()) {
        const uint8_t *src = dT + segs[0].offset;
        uint8_t *dst = dU + segs[0].offset;
        if (!direct) {
            grow(w.T, (size_t)N + 64);
            for (size_t k = 0; k < bl.size(); k++)
                HIPCHK(hipMemcpyAsync(P<uint8_t>(w.T) + bl[k].off, dT + segs[which[k]].offset, bl[k].n,
                                      hipMemcpyDeviceToDevice, s));
            src = P<uint8_t>(w.T);
            dst = P<uint8_t>(w.T);
        }
        inverse(w, src, dst, bl, N, s);
        if (!direct)
            for (size_t k = 0; k < bl.size(); k++)
                HIPCHK(hipMemcpyAsync(dU + segs[which[k]].offset, P<uint8_t>(w.T) + bl[k].off, bl[k].n,
                                      hipMemcpyDeviceToDevice, s));
    }
    for (int32_t b = 0; b < count; b++)
        if (segs[b].n == 1)
            HIPCHK(hipMemcpyAsync(dU + segs[b].offset, dT + segs[b].offset, 1, hipMemcpyDeviceToDevice, s));
    SYNCCHK(s);
}

bool samples_ok(const kbwt_seg *segs, int32_t count, const int32_t *I) {
    for (int32_t b = 0, k = 0; b < count; k += (int32_t)nsamples(segs[b].n, segs[b].r), b++) {
        const int32_t n = segs[b].n;
        if (n <= 1) {
            if (I[k] != n) return false;
            continue;
        }
        for (int64_t t = 0; t <= (n - 1) / segs[b].r; ++t)
            if (I[k + t] <= 0 || I[k + t] > n) return false;
    }
    return true;
}

// After an error inside a call: leave capture mode if the error came during a capture, let every launch queued so far
// finish (each is a whole, valid launch, so each ends), drop the cached inverse graph (its buffers may have been
// freed by a failed allocation) and the profiling events of unfinished launches, and clear the runtime's last error.
// The next call then starts from a clean workspace. Never throws.
void recover(Workspace &w, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    }
    (void)hipStreamSynchronize(s);
    if (s != w.stream) (void)hipStreamSynchronize(w.stream);
    for (int b = 0; b < 2; b++)
        if (w.pins.ev[b]) (void)hipEventSynchronize(w.pins.ev[b]);
    w.inv_graph.clear();
    w.fwd_graph.clear();
    try {
        for (auto &p : w.pending) {
            w.spare.push_back(p.a);
            w.spare.push_back(p.b);
        }
    } catch (...) {
    }
    w.pending.clear();
    (void)hipGetLastError();
}

}  // namespace

void record_error(const char *msg) {
    g_last_error = msg;
    fprintf(stderr, "kbwt: %s\n", msg);
}

}  // namespace kbwt

using namespace kbwt;

template <class F> static int32_t guarded(F f) {
    g_last_error.clear();
    try {
        f();
        return 0;
    } catch (const DevError &e) {
        g_counters.device_errors++;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return KBWT_ERR_NOMEM;
    }
}

// f(w, s) on the current device's workspace, which it holds locked; s is the caller's stream, or the library's. On
// an error the workspace recovers (recover) before the code goes back to the caller, so the next call is unaffected.
template <class F> static int32_t on_device(void *stream, F f) {
    return guarded([&] {
        Workspace &w = workspace();
        std::lock_guard<std::mutex> lk(w.mu);
        hipStream_t s = stream ? (hipStream_t)stream : w.stream;
        try {
            f(w, s);
        } catch (...) {
            recover(w, s);
            throw;
        }
    });
}

// The CPU path after a device error: off unless the caller enables it (kbwt_set_cpu_fallback, or KBWT_CPU_FALLBACK=1
// in the environment when the library first looks).
static bool fallback_on() {
    int v = g_cpu_fallback.load();
    if (v < 0) {
        const char *e = getenv("KBWT_CPU_FALLBACK");
        v = e && e[0] == '1' ? 1 : 0;
        int expect = -1;
        if (!g_cpu_fallback.compare_exchange_strong(expect, v)) v = expect;
    }
    return v == 1;
}

// One block through the CPU path after its device work failed with rc (-2 or -3). Returns the block's new code.
static int32_t cpu_fallback(bool fwd, const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *Iout,
                            const int32_t *Iin, int32_t threads, int32_t rc) {
    if (rc == KBWT_OK || rc == KBWT_ERR_ARG || !fallback_on()) return rc;
    const int32_t rc2 = fwd ? cpu_bwt_aux(T, U, n, r, Iout, threads) : cpu_unbwt_aux(T, U, n, r, Iin, threads);
    if (rc2 != KBWT_OK) return rc;
    if (g_counters.cpu_fallbacks++ == 0)
        fprintf(stderr, "kbwt: a device call failed; its blocks run on the CPU path (KBWT_CPU_FALLBACK), counted by "
                        "kbwt_get_counters\n");
    return KBWT_OK;
}

// ---- concurrent single-block calls run as batches (kbwt_combine.h): the runner of one batch of one kind, from host
// buffers packed back to back in the workspace's staging buffers
static Combiner g_combiner;

// A lane's stream and events, made on its first batch.
static void lane_ready(Lane &ln) {
    if (ln.ready) return;
    if (!ln.cs) HIPCHK(hipStreamCreateWithFlags(&ln.cs, hipStreamNonBlocking));
    if (!ln.staged) HIPCHK(hipEventCreateWithFlags(&ln.staged, hipEventDisableTiming));
    if (!ln.computed) HIPCHK(hipEventCreateWithFlags(&ln.computed, hipEventDisableTiming));
    ln.ready = true;
}

// After an error in a lane's batch: let its queued copies finish and clear the runtime's last error. Never throws.
static void recover_lane(Lane &ln) {
    if (ln.cs) (void)hipStreamSynchronize(ln.cs);
    for (int b = 0; b < 2; b++)
        if (ln.pins.ev[b]) (void)hipEventSynchronize(ln.pins.ev[b]);
    (void)hipGetLastError();
}

// One combined batch on lane `lane`, in three steps: the callers' bytes to the lane's device buffer on the lane's copy
// stream (no workspace lock, so another batch's kernels run meanwhile), then in.done() and the kernels on the
// workspace's stream once the copy is in (workspace lock), then the results back to the callers' buffers on the
// lane's stream, while the next batch's kernels run.
static void run_pending_batch(std::vector<Pending *> &b, bool fwd, int lane, Intake &in) {
    const int32_t count = (int32_t)b.size();
    std::vector<kbwt_seg> segs(count);
    int64_t at = 0;
    size_t ns = 0;
    for (int32_t k = 0; k < count; k++) {
        segs[k].offset = at;
        segs[k].n = b[k]->n;
        segs[k].r = b[k]->r;
        at += b[k]->n;
        ns += nsamples(b[k]->n, b[k]->r);
    }
    std::vector<int32_t> I(ns);
    if (!fwd)
        for (int32_t k = 0, o = 0; k < count; o += (int32_t)nsamples(b[k]->n, b[k]->r), k++)
            memcpy(I.data() + o, b[k]->Iin, nsamples(b[k]->n, b[k]->r) * 4);
    std::vector<Span> sp(count);
    for (int32_t k = 0; k < count; k++) sp[k] = {const_cast<uint8_t *>(b[k]->T), (size_t)b[k]->n};
    Lane *lp = nullptr;
    int32_t rc = guarded([&] {
        Workspace &w = workspace();
        lp = &w.lane[lane];
    });
    std::unique_lock<std::mutex> lane_lk;
    if (rc == 0) lane_lk = std::unique_lock<std::mutex>(lp->mu);
    uint8_t *out = nullptr;
    if (rc == 0)
        rc = guarded([&] {
            Lane &ln = *lp;
            try {
                lane_ready(ln);
                grow(ln.T, (size_t)at + 64);
                if (fwd) grow(ln.U, (size_t)at + 64);
                out = fwd ? P<uint8_t>(ln.U) : P<uint8_t>(ln.T);
                to_device(ln.pins, P<uint8_t>(ln.T), sp, ln.cs);
                HIPCHK(hipEventRecord(ln.staged, ln.cs));
            } catch (...) {
                recover_lane(ln);
                throw;
            }
        });
    in.done();
    if (rc == 0)
        rc = on_device(nullptr, [&](Workspace &w, hipStream_t s) {
            // no graph capture here: combined batches seldom repeat a shape, and other callers' threads allocate
            // and copy meanwhile (their lanes' staging)
            GraphsOff off;
            HIPCHK(hipStreamWaitEvent(s, lp->staged, 0));
            if (fwd) forward_batch_dev(w, P<uint8_t>(lp->T), out, segs.data(), count, I.data(), s);
            else inverse_batch_dev(w, P<uint8_t>(lp->T), out, segs.data(), count, I.data(), s);
            HIPCHK(hipEventRecord(lp->computed, s));
        });
    if (rc == 0)
        rc = guarded([&] {
            Lane &ln = *lp;
            try {
                HIPCHK(hipStreamWaitEvent(ln.cs, ln.computed, 0));
                for (int32_t k = 0; k < count; k++) sp[k].p = b[k]->U;
                to_host(ln.pins, sp, out, ln.cs);
            } catch (...) {
                recover_lane(ln);
                throw;
            }
        });
    if (rc != 0 && lp) recover_lane(*lp);
    const std::string err = g_last_error;
    for (int32_t k = 0, o = 0; k < count; o += (int32_t)nsamples(b[k]->n, b[k]->r), k++) {
        b[k]->rc = rc;
        b[k]->err = err;
        if (fwd && rc == 0) memcpy(b[k]->I, I.data() + o, nsamples(b[k]->n, b[k]->r) * 4);
    }
}

// Never throws: a caller that runs a batch must always wake the others. A block that fails on its own (the combiner
// runs the blocks of a failed batch again one by one) takes the CPU path when it is enabled.
static void run_pending(std::vector<Pending *> &b, bool fwd, int lane, Intake &in) {
    if (b.empty()) return;
    try {
        run_pending_batch(b, fwd, lane, in);
    } catch (...) {
        for (Pending *q : b) {
            q->rc = KBWT_ERR_NOMEM;
            q->err = "host allocation failed";
        }
    }
    if (b.size() == 1 && b[0]->rc != KBWT_OK) {
        Pending *q = b[0];
        if (cpu_fallback(fwd, q->T, q->U, q->n, q->r, q->I, q->Iin, q->threads, q->rc) == KBWT_OK) {
            q->rc = KBWT_OK;
            q->err.clear();
        }
    }
}

static int32_t combine(Pending &p) {
    std::string err;
    const int32_t rc = g_combiner.submit(p, run_pending, err);
    g_last_error = err;
    return rc;
}

// A workspace's device buffers, cached graphs and pinned chunks, freed (the caller holds its lock). Never throws.
static void release_buffers(Workspace &w) {
    Buf *bufs[] = {&w.T,       &w.U,      &w.K0,    &w.K1,     &w.V0,     &w.V1,     &w.SA,     &w.ISA,    &w.L,
                   &w.pos[0],  &w.pos[1], &w.grp[0], &w.grp[1], &w.status, &w.st_sum, &w.carry, &w.samples,
                   &w.stage,   &w.blk,    &w.counts, &w.dbase,  &w.tot,    &w.cbase,  &w.bad,    &w.meta,
                   &w.gh[0],   &w.gh[1],  &w.gs[0],  &w.gs[1],  &w.lh[0],  &w.lh[1],  &w.ls[0],  &w.ls[1],
                   &w.go,      &w.sec,    &w.lo,     &w.tstart, &w.act,    &w.rfeat,  &w.bigram, &w.rmeta,
                   &w.sufl,    &w.eh,     &w.es,     &w.ecnt,   &w.eoff,   &w.pair,   &w.tmax,   &w.lrow};
    w.inv_graph.clear();
    w.fwd_graph.clear();
    for (Buf *b : bufs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->cap = 0;
    }
    release_pins(w.pins);
}

extern "C" {

// A block longer than one batch may hold (BATCH_MAX_TOTAL, 2^31 - 65 bytes) cannot be indexed by the device
// kernels: the drop-in entry points report it as out of memory (-2) before anything is queued.
static int32_t too_long(int32_t n) {
    if ((int64_t)n <= BATCH_MAX_TOTAL) return 0;
    g_last_error = "block longer than 2^31 - 65 bytes";
    return KBWT_ERR_NOMEM;
}

int32_t kolmo_sais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, int32_t *I,
                           int32_t threads) {
    if (T == NULL || U == NULL || A == NULL || n < 0 || r < 2 || (r & (r - 1)) != 0 || I == NULL) return -1;
    if (n <= 1) {
        if (n == 1) U[0] = T[0];
        I[0] = n;
        return 0;
    }
    if (too_long(n)) return KBWT_ERR_NOMEM;
    Pending p;
    p.threads = threads;
    p.fwd = true;
    p.T = T;
    p.U = U;
    p.n = n;
    p.r = r;
    p.I = I;
    p.Iin = nullptr;
    return combine(p);
}

int32_t kolmo_sais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, const int32_t *I,
                             int32_t threads) {
    if (T == NULL || U == NULL || A == NULL || n < 0 || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) || I == NULL)
        return -1;
    if (n <= 1) {
        if (I[0] != n) return -1;
        if (n == 1) U[0] = T[0];
        return 0;
    }
    for (int64_t t = 0; t <= (n - 1) / r; ++t)
        if (I[t] <= 0 || I[t] > n) return -1;
    if (too_long(n)) return KBWT_ERR_NOMEM;
    Pending p;
    p.threads = threads;
    p.fwd = false;
    p.T = T;
    p.U = U;
    p.n = n;
    p.r = r;
    p.I = nullptr;
    p.Iin = I;
    return combine(p);
}

int32_t kolmo_sais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t i, int32_t threads) {
    return kolmo_sais_unbwt_aux(T, U, A, n, n, &i, threads);
}

int32_t kbwt_dev_bwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, int32_t *I, void *stream) {
    if (dT == NULL || dU == NULL || !forward_args_ok(n, r) || I == NULL) return -1;
    if (too_long(n)) return KBWT_ERR_NOMEM;
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        kbwt_seg g{0, n, r};
        forward_batch_dev(w, dT, dU, &g, 1, I, s);
    });
}

int32_t kbwt_dev_unbwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, const int32_t *I, void *stream) {
    if (dT == NULL || dU == NULL || !inverse_args_ok(n, r) || I == NULL) return -1;
    kbwt_seg g{0, n, r};
    if (!samples_ok(&g, 1, I)) return -1;
    if (too_long(n)) return KBWT_ERR_NOMEM;
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        inverse_batch_dev(w, dT, dU, &g, 1, I, s);
    });
}

int32_t kbwt_dev_unbwt_aux_noalloc(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, const int32_t *I,
                                   void *stream) {
    if (dT == NULL || dU == NULL || !inverse_args_ok(n, r) || I == NULL) return -1;
    kbwt_seg g{0, n, r};
    if (!samples_ok(&g, 1, I)) return -1;
    if ((int64_t)n > (int64_t)NOALLOC_MAX_N) {
        g_last_error = "kbwt_dev_unbwt_aux_noalloc: n exceeds its scratch limit of 4194304 bytes";
        return KBWT_ERR_ARG;
    }
    return guarded([&] {
        hipStream_t s = (hipStream_t)stream;
        try {
            inverse_noalloc(dT, dU, (uint32_t)n, (uint32_t)I[0], s);
        } catch (...) {   // let what was queued finish before the error goes back
            (void)hipStreamSynchronize(s);
            (void)hipGetLastError();
            throw;
        }
    });
}

int32_t kbwt_dev_bwt_aux_batch_features(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                                        int32_t *I, float *features, void *stream) {
    if (dT == NULL || dU == NULL || I == NULL || features == NULL || !batch_ok(segs, count, true)) return -1;
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        forward_batch_dev(w, dT, dU, segs, count, I, s, features);
    });
}
int32_t kbwt_dev_bwt_aux_batch(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count, int32_t *I,
                               void *stream) {
    if (dT == NULL || dU == NULL || I == NULL || !batch_ok(segs, count, true)) return -1;
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        forward_batch_dev(w, dT, dU, segs, count, I, s);
    });
}

int32_t kbwt_dev_unbwt_aux_batch(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                                 const int32_t *I, void *stream) {
    if (dT == NULL || dU == NULL || I == NULL || !batch_ok(segs, count, false)) return -1;
    if (!samples_ok(segs, count, I)) return -1;
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        inverse_batch_dev(w, dT, dU, segs, count, I, s);
    });
}

// ---- autotuner: coordinate descent over the launch parameters, timed on the caller's own blocks

static bool tuning_ok(const kbwt_tuning *t) {
    return t && t->sort_cfg >= 0 && t->sort_cfg < SORT_NCFG && t->inv_dshift >= 2 && t->inv_dshift <= 12 &&
           t->seg_rank >= 0 && t->seg_rank <= 1024;
}

int32_t kbwt_set_tuning(int32_t size_log2, const kbwt_tuning *t) {
    if ((t != NULL && !tuning_ok(t)) || size_log2 >= TUNE_CLASSES) return KBWT_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int c = 0; c < TUNE_CLASSES; c++)
        if (size_log2 < 0 || c == size_log2) {
            if (t == NULL) {   // back to the built-in table
                g_tune[c] = default_tuning((uint32_t)c);
                continue;
            }
            g_tune[c].sort_cfg = (uint32_t)t->sort_cfg;
            g_tune[c].inv_dshift = (uint32_t)t->inv_dshift;
            g_tune[c].seg_rank = (uint32_t)t->seg_rank;
        }
    return 0;
}

int32_t kbwt_get_tuning(int32_t size_log2, kbwt_tuning *t) {
    if (t == NULL || size_log2 < 0 || size_log2 >= TUNE_CLASSES) return KBWT_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    t->sort_cfg = (int32_t)g_tune[size_log2].sort_cfg;
    t->inv_dshift = (int32_t)g_tune[size_log2].inv_dshift;
    t->seg_rank = (int32_t)g_tune[size_log2].seg_rank;
    t->ms = 0;
    return 0;
}

int32_t kbwt_autotune(const uint8_t *dT, const kbwt_seg *segs, int32_t count, void *stream, kbwt_tuning *best,
                      kbwt_tuning *start) {
    if (dT == NULL || best == NULL || !batch_ok(segs, count, true) || count == 0) return -1;
    int64_t total = 0;
    for (int32_t b = 0; b < count; b++) total += segs[b].n;
    if (total < 2) return -1;
    const uint32_t cls = size_class((uint32_t)total);
    return on_device(stream, [&](Workspace &w, hipStream_t s) {
        // the class's entry as it was, put back unless the search finishes (a trial that fails leaves nothing
        // untested installed); the scratch outputs and events are released on every path
        struct Restore {
            uint32_t cls;
            Tuning saved;
            bool keep = false;
            ~Restore() {
                if (keep) return;
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tune[cls] = saved;
            }
        } restore{cls, tuning_for_class(cls)};
        struct DevBuf {
            void *p = nullptr;
            ~DevBuf() {
                if (p) (void)hipFree(p);
            }
        } bU, bB;
        struct Event {
            hipEvent_t e = nullptr;
            ~Event() {
                if (e) (void)hipEventDestroy(e);
            }
        } ev0, ev1;
        const int64_t span = segs[count - 1].offset + segs[count - 1].n;
        fault_point(FS_ALLOC);
        HIPCHK(hipMalloc(&bU.p, (size_t)span + 64));
        HIPCHK(hipMalloc(&bB.p, (size_t)span + 64));
        void *dU = bU.p, *dB = bB.p;
        std::vector<int32_t> I;
        size_t ns = 0;
        for (int32_t b = 0; b < count; b++) ns += nsamples(segs[b].n, segs[b].r);
        I.resize(ns);
        HIPCHK(hipEventCreate(&ev0.e));
        HIPCHK(hipEventCreate(&ev1.e));
        hipEvent_t e0 = ev0.e, e1 = ev1.e;
        // one forward + inverse with tuning t installed for this size class; best of two timed runs after a warm-up
        auto trial = [&](const Tuning &t) {
            {
                std::lock_guard<std::mutex> tl(g_tune_mu);
                g_tune[cls] = t;
            }
            float best_ms = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                HIPCHK(hipEventRecord(e0, s));
                forward_batch_dev(w, dT, (uint8_t *)dU, segs, count, I.data(), s);
                inverse_batch_dev(w, (const uint8_t *)dU, (uint8_t *)dB, segs, count, I.data(), s);
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0 && ms < best_ms) best_ms = ms;
            }
            return best_ms;
        };
        Tuning cur = default_tuning(cls);
        if (start && tuning_ok(start)) {
            cur.sort_cfg = (uint32_t)start->sort_cfg;
            cur.inv_dshift = (uint32_t)start->inv_dshift;
            cur.seg_rank = (uint32_t)start->seg_rank;
        }
        float cur_ms = trial(cur);
        const float start_ms = cur_ms;
        const uint32_t dshifts[] = {3, 4, 5, 6, 7}, ranks[] = {0, 16, 32, 64, 128};
        for (int pass = 0; pass < 2; pass++) {
            bool moved = false;
            for (uint32_t c = 0; c < (uint32_t)SORT_NCFG; c++) {
                if (c == cur.sort_cfg) continue;
                Tuning t = cur;
                t.sort_cfg = c;
                const float ms = trial(t);
                if (ms < cur_ms) { cur = t; cur_ms = ms; moved = true; }
            }
            for (uint32_t d : dshifts) {
                if (d == cur.inv_dshift) continue;
                Tuning t = cur;
                t.inv_dshift = d;
                const float ms = trial(t);
                if (ms < cur_ms) { cur = t; cur_ms = ms; moved = true; }
            }
            for (uint32_t r : ranks) {
                if (r == cur.seg_rank) continue;
                Tuning t = cur;
                t.seg_rank = r;
                const float ms = trial(t);
                if (ms < cur_ms) { cur = t; cur_ms = ms; moved = true; }
            }
            if (!moved) break;
        }
        {
            std::lock_guard<std::mutex> tl(g_tune_mu);
            g_tune[cls] = cur;
            restore.keep = true;
        }
        best->sort_cfg = (int32_t)cur.sort_cfg;
        best->inv_dshift = (int32_t)cur.inv_dshift;
        best->seg_rank = (int32_t)cur.seg_rank;
        best->ms = cur_ms;
        if (start) start->ms = start_ms;
    });
}

// Host-pointer batches: the span [first offset, last end) goes to the device once, keeping the offsets. When the
// device work fails and the CPU path is enabled, every block of the batch is computed there instead.
struct DevSpan {   // a device buffer freed on every path out of a call
    void *p = nullptr;
    ~DevSpan() {
        if (p) (void)hipFree(p);
    }
};

int32_t kbwt_bwt_aux_batch(const uint8_t *T, uint8_t *U, const kbwt_seg *segs, int32_t count, int32_t *I) {
    if (T == NULL || U == NULL || I == NULL || !batch_ok(segs, count, true)) return -1;
    if (count == 0) return 0;
    const int32_t rc = on_device(nullptr, [&](Workspace &w, hipStream_t s) {
        const int64_t lo = segs[0].offset, hi = segs[count - 1].offset + segs[count - 1].n;
        DevSpan in, out;
        fault_point(FS_ALLOC);
        HIPCHK(hipMalloc(&in.p, (size_t)(hi - lo) + 64));
        HIPCHK(hipMalloc(&out.p, (size_t)(hi - lo) + 64));
        to_device(w.pins, (uint8_t *)in.p, {{const_cast<uint8_t *>(T) + lo, (size_t)(hi - lo)}}, s);
        std::vector<kbwt_seg> g(segs, segs + count);
        for (auto &x : g) x.offset -= lo;
        forward_batch_dev(w, (const uint8_t *)in.p, (uint8_t *)out.p, g.data(), count, I, s);
        to_host(w.pins, {{U + lo, (size_t)(hi - lo)}}, (const uint8_t *)out.p, s);
        SYNCCHK(s);
    });
    if (rc == KBWT_OK || !fallback_on()) return rc;
    for (int32_t b = 0, k = 0; b < count; k += (int32_t)nsamples(segs[b].n, segs[b].r), b++)
        if (cpu_fallback(true, T + segs[b].offset, U + segs[b].offset, segs[b].n, segs[b].r, I + k, nullptr, 1, rc) !=
            KBWT_OK)
            return rc;
    g_last_error.clear();
    return KBWT_OK;
}

int32_t kbwt_unbwt_aux_batch(const uint8_t *T, uint8_t *U, const kbwt_seg *segs, int32_t count, const int32_t *I) {
    if (T == NULL || U == NULL || I == NULL || !batch_ok(segs, count, false)) return -1;
    if (!samples_ok(segs, count, I)) return -1;
    if (count == 0) return 0;
    const int32_t rc = on_device(nullptr, [&](Workspace &w, hipStream_t s) {
        const int64_t lo = segs[0].offset, hi = segs[count - 1].offset + segs[count - 1].n;
        DevSpan io;
        fault_point(FS_ALLOC);
        HIPCHK(hipMalloc(&io.p, (size_t)(hi - lo) + 64));
        to_device(w.pins, (uint8_t *)io.p, {{const_cast<uint8_t *>(T) + lo, (size_t)(hi - lo)}}, s);
        std::vector<kbwt_seg> g(segs, segs + count);
        for (auto &x : g) x.offset -= lo;
        inverse_batch_dev(w, (const uint8_t *)io.p, (uint8_t *)io.p, g.data(), count, I, s);
        to_host(w.pins, {{U + lo, (size_t)(hi - lo)}}, (const uint8_t *)io.p, s);
        SYNCCHK(s);
    });
    if (rc == KBWT_OK || !fallback_on()) return rc;
    for (int32_t b = 0, k = 0; b < count; k += (int32_t)nsamples(segs[b].n, segs[b].r), b++)
        if (cpu_fallback(false, T + segs[b].offset, U + segs[b].offset, segs[b].n, segs[b].r, nullptr, I + k, 1, rc) !=
            KBWT_OK)
            return rc;
    g_last_error.clear();
    return KBWT_OK;
}

int32_t kbwt_set_cpu_fallback(int32_t on) {
    const bool was = fallback_on();
    g_cpu_fallback.store(on ? 1 : 0);
    return was ? 1 : 0;
}

int32_t kbwt_fault_inject(const char *spec) {
    int site;
    uint64_t at, count;
    if (!parse_fault(spec, site, at, count)) return KBWT_ERR_ARG;
    arm_fault(site, at, count);
    return KBWT_OK;
}

int32_t kbwt_get_counters(kbwt_counters *out) {
    if (!out) return KBWT_ERR_ARG;
    uint64_t batches = 0, blocks = 0, retried = 0;
    g_combiner.counts(batches, blocks, retried);
    out->device_errors = g_counters.device_errors.load();
    out->retried_blocks = (int64_t)retried;
    out->cpu_fallbacks = g_counters.cpu_fallbacks.load();
    out->injected_faults = g_counters.injected_faults.load();
    out->combined_batches = (int64_t)batches;
    out->combined_blocks = (int64_t)blocks;
    out->graph_replays = g_counters.graph_replays.load();
    return KBWT_OK;
}

void kbwt_profile(int32_t on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
    g_prof_only.clear();
    g_prof.clear();
}

void kbwt_profile_only(const char *name) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = name != nullptr && name[0] != 0;
    g_prof_only = g_prof_on ? name : "";
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
        SYNCCHK(s);
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
    // the lanes first (a batch holds its lane, then the workspace), so that no batch in flight loses its buffers
    std::unique_lock<std::mutex> lanes[COMBINE_LANES];
    for (int k = 0; k < COMBINE_LANES; k++) lanes[k] = std::unique_lock<std::mutex>(w.lane[k].mu);
    std::lock_guard<std::mutex> lk(w.mu);
    for (Lane &ln : w.lane) {
        if (ln.cs) (void)hipStreamSynchronize(ln.cs);
        for (Buf *b : {&ln.T, &ln.U}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
        release_pins(ln.pins);
    }
    release_buffers(w);
}

}  // extern "C"
