/* Seeded synthetic blocks for the block-sorting path: the five measured workloads and the input kinds of kolmo's own
 * block-sorting tests. Host C, no GPU. Every generator is a pure function of (seed, n): the same bytes on any machine.
 *
 * PRNG: xoshiro256** (Blackman, Vigna 2018) seeded through splitmix64. Normals: Box-Muller. Gamma(a): Marsaglia-Tsang
 * (2000), with Gamma(a) = Gamma(a + 1) * U^(1/a) for a < 1. Dirichlet(a) rows: normalised Gamma(a) draws.
 *
 * Workloads (sizes are the caller's n):
 *   kbwt_gen_markov2  order-2 Markov text over the 64 symbols 0x21..0x60, one Dirichlet(0.3) row per context.
 *   kbwt_gen_mixed    lines: order-3 Markov text over '\n' + 0x20..0x7E (96 symbols, Dirichlet(0.2) rows, lines of at
 *                     most 200 bytes) interleaved with CSV lines of 8 int32 random-walk columns (N(0,50) steps), so that
 *                     CSV lines make 30% of the bytes.
 *   kbwt_gen_repeats  1 MiB of uniform bytes, then copies of earlier substrings (length log-uniform 64 B..4 MiB, source
 *                     uniform over earlier data, overlap allowed) with 0.01% substituted bytes, and one 2 MiB run of a
 *                     single byte placed uniformly.
 *   kbwt_gen_dna      A/C/G/T with 1% N; at each base a tandem repeat starts with probability 1/1000 (unit 2..50 bases,
 *                     5..200 copies).
 *   kbwt_gen_random   uniform bytes.
 *   kbwt_gen_csvcols  CSV lines of 8 decimal random-walk columns only.
 *   kbwt_gen_logs     template log lines with an increasing integer timestamp, an IPv4 address and small integers.
 *   kbwt_gen_bf16     bf16 weights drawn from N(0, 0.02), little-endian, round to nearest even.
 *   kbwt_gen_bsc_kind the nine input kinds of kolmo's tests/test_bwt.c:24-45 (gen()), restated byte for byte: the
 *                     golden hashes of tests/test_bwt_golden.h are over these inputs.
 * Batches: kbwt_gen_batch_plan draws block sizes log-uniform in [lo, hi] and a generator per block. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static inline double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint64_t rng_below(rng_t *r, uint64_t k) { return k ? rng_next(r) % k : 0; }
static double rng_normal(rng_t *r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static double rng_gamma(rng_t *r, double a) {
    if (a < 1.0) {
        double u = rng_unif(r);
        if (u < 1e-300) u = 1e-300;
        return rng_gamma(r, a + 1.0) * pow(u, 1.0 / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x, v;
        do { x = rng_normal(r); v = 1.0 + c * x; } while (v <= 0.0);
        v = v * v * v;
        double u = rng_unif(r);
        if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
    }
}
/* A Dirichlet(a) row over k <= 256 symbols as a Walker/Vose alias table: entry i packs a 24-bit acceptance threshold
   (high bits) and the alias symbol (low 8 bits). Sampling reads one entry: one cache line per draw. */
static void dirichlet_alias(uint64_t seed, int k, double a, uint32_t *tab) {
    rng_t r; rng_seed(&r, seed);
    double p[256], s = 0;
    int small[256], large[256], ns = 0, nl = 0;
    for (int i = 0; i < k; i++) { p[i] = rng_gamma(&r, a); s += p[i]; }
    if (s <= 0) { for (int i = 0; i < k; i++) p[i] = 1.0; s = k; }
    for (int i = 0; i < k; i++) {
        p[i] = p[i] / s * k;
        if (p[i] < 1.0) small[ns++] = i; else large[nl++] = i;
    }
    while (ns && nl) {
        int l = small[--ns], g = large[--nl];
        double t = p[l] * 16777216.0;
        tab[l] = ((t >= 16777215.0 ? 16777215u : (uint32_t)t) << 8) | (uint32_t)g;
        p[g] = p[g] + p[l] - 1.0;
        if (p[g] < 1.0) small[ns++] = g; else large[nl++] = g;
    }
    while (nl) { int g = large[--nl]; tab[g] = (0xFFFFFFu << 8) | (uint32_t)g; }
    while (ns) { int l = small[--ns]; tab[l] = (0xFFFFFFu << 8) | (uint32_t)l; }
}
static inline int alias_pick(const uint32_t *tab, int k, uint64_t u) {
    int i = (int)(((u >> 32) * (uint64_t)k) >> 32);
    uint32_t e = tab[i];
    return ((uint32_t)u & 0xFFFFFFu) < (e >> 8) ? i : (int)(e & 0xFFu);
}
static uint64_t row_seed(uint64_t seed, uint64_t ctx) {
    uint64_t x = seed * 0xD1B54A32D192ED03ull ^ (ctx + 1) * 0x9E3779B97F4A7C15ull;
    return splitmix64(&x);
}

/* ------------------------------------------------------------------ order-2 Markov, 64 symbols */

int kbwt_gen_markov2(uint64_t model_seed, uint64_t seed, uint8_t *out, size_t n) {
    enum { K = 64 };
    uint32_t *cdf = (uint32_t *)malloc((size_t)K * K * K * sizeof(uint32_t));   /* 4096 alias rows of 64 */
    if (!cdf) return -2;
    for (int c = 0; c < K * K; c++) dirichlet_alias(row_seed(model_seed, (uint64_t)c), K, 0.3, cdf + (size_t)c * K);
    rng_t r; rng_seed(&r, seed);
    int c1 = 0, c2 = 0;
    for (size_t i = 0; i < n; i++) {
        int s = alias_pick(cdf + (size_t)(c1 * K + c2) * K, K, rng_next(&r));
        out[i] = (uint8_t)(0x21 + s);
        c1 = c2; c2 = s;
    }
    free(cdf);
    return 0;
}

/* ------------------------------------------------------------------ order-3 Markov text + CSV lines */

/* The order-3 model: one alias row per context (96^3 contexts), all drawn up front on every core, cached by model
   seed for the next call (the rows are a pure function of (model seed, context)). */
enum { M3K = 96, M3CTX = 96 * 96 * 96 };
typedef struct {
    uint64_t seed;
    uint32_t *rows;         /* M3CTX alias tables of 96 entries */
} markov3_t;
static pthread_mutex_t m3_mu = PTHREAD_MUTEX_INITIALIZER;
static markov3_t m3_cache = {0, NULL};
static const markov3_t *m3_get(uint64_t seed) {
    pthread_mutex_lock(&m3_mu);
    if (!m3_cache.rows || m3_cache.seed != seed) {
        if (!m3_cache.rows) m3_cache.rows = (uint32_t *)malloc((size_t)M3CTX * M3K * sizeof(uint32_t));
        if (!m3_cache.rows) { pthread_mutex_unlock(&m3_mu); return NULL; }
        uint32_t *rows = m3_cache.rows;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < M3CTX; c++) dirichlet_alias(row_seed(seed, (uint64_t)c), M3K, 0.2, rows + (size_t)c * M3K);
        m3_cache.seed = seed;
    }
    pthread_mutex_unlock(&m3_mu);
    return &m3_cache;
}
static inline const uint32_t *m3_row(const markov3_t *m, int ctx) { return m->rows + (size_t)ctx * M3K; }
static inline uint8_t m3_byte(int s) { return s == 0 ? (uint8_t)'\n' : (uint8_t)(0x1F + s); }

/* one text line (<= 200 bytes, ends with '\n') into buf; returns its length */
static int m3_line(const markov3_t *m, rng_t *r, int *ctx, uint8_t *buf) {
    int len = 0;
    for (;;) {
        const uint32_t *row = m3_row(m, *ctx);
        int s = len == 199 ? 0 : alias_pick(row, M3K, rng_next(r));
        buf[len++] = m3_byte(s);
        *ctx = (*ctx % (M3K * M3K)) * M3K + s;
        if (s == 0) return len;
    }
}
static int csv_line(rng_t *r, int64_t *cols, char *buf) {
    int len = 0;
    for (int c = 0; c < 8; c++) {
        cols[c] += (int64_t)llround(rng_normal(r) * 50.0);
        if (cols[c] > INT32_MAX) cols[c] = INT32_MAX;
        if (cols[c] < INT32_MIN) cols[c] = INT32_MIN;
        len += snprintf(buf + len, 32, c ? ",%d" : "%d", (int)cols[c]);
    }
    buf[len++] = '\n';
    return len;
}

int kbwt_gen_mixed2(uint64_t model_seed, uint64_t seed, uint8_t *out, size_t n) {
    const markov3_t *m = m3_get(model_seed);
    if (!m) return -2;
    rng_t r; rng_seed(&r, seed);
    rng_t rc; rng_seed(&rc, seed ^ 0xC5C5C5C5C5C5C5C5ull);
    int ctx = 0;
    int64_t cols[8] = {0};
    size_t total = 0, csv = 0;
    uint8_t line[256];
    while (total < n) {
        int len;
        int is_csv = csv * 10 < total * 3;
        if (is_csv) len = csv_line(&rc, cols, (char *)line);
        else len = m3_line(m, &r, &ctx, line);
        size_t k = (size_t)len < n - total ? (size_t)len : n - total;
        memcpy(out + total, line, k);
        total += k;
        if (is_csv) csv += k;
    }
    return 0;
}
int kbwt_gen_mixed(uint64_t seed, uint8_t *out, size_t n) { return kbwt_gen_mixed2(seed, seed, out, n); }

/* ------------------------------------------------------------------ deep repeats */

int kbwt_gen_repeats(uint64_t seed, uint8_t *out, size_t n) {
    rng_t r; rng_seed(&r, seed);
    const size_t head = n < ((size_t)1 << 20) ? n : ((size_t)1 << 20);
    for (size_t i = 0; i < head; i++) out[i] = (uint8_t)(rng_next(&r) >> 56);
    const size_t run = (size_t)2 << 20;
    size_t run_at = n, run_len = 0;
    if (n >= head + run) { run_at = head + rng_below(&r, n - head - run + 1); run_len = run; }
    const uint8_t run_byte = (uint8_t)(rng_next(&r) >> 56);
    size_t i = head;
    const double lmin = log(64.0), lmax = log(4194304.0);
    while (i < n) {
        if (i == run_at) { memset(out + i, run_byte, run_len); i += run_len; continue; }
        size_t L = (size_t)exp(lmin + (lmax - lmin) * rng_unif(&r));
        if (L > n - i) L = n - i;
        if (i < run_at && L > run_at - i) L = run_at - i;
        size_t src = rng_below(&r, i);
        for (size_t k = 0; k < L; k++) {
            uint8_t b = out[src + k];
            if (rng_next(&r) % 10000 == 0) b = (uint8_t)(rng_next(&r) >> 56);
            out[i + k] = b;
        }
        i += L;
    }
    return 0;
}

/* ------------------------------------------------------------------ DNA, random, CSV columns, logs, bf16 */

int kbwt_gen_dna(uint64_t seed, uint8_t *out, size_t n) {
    static const char B[4] = {'A', 'C', 'G', 'T'};
    rng_t r; rng_seed(&r, seed);
    size_t i = 0;
    while (i < n) {
        if (rng_next(&r) % 1000 == 0) {
            int unit = 2 + (int)rng_below(&r, 49), copies = 5 + (int)rng_below(&r, 196);
            char u[50];
            for (int k = 0; k < unit; k++) u[k] = B[rng_next(&r) >> 62];
            for (int c = 0; c < copies && i < n; c++)
                for (int k = 0; k < unit && i < n; k++) out[i++] = (uint8_t)u[k];
            continue;
        }
        out[i++] = rng_next(&r) % 100 == 0 ? (uint8_t)'N' : (uint8_t)B[rng_next(&r) >> 62];
    }
    return 0;
}

int kbwt_gen_random(uint64_t seed, uint8_t *out, size_t n) {
    rng_t r; rng_seed(&r, seed);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) { uint64_t x = rng_next(&r); memcpy(out + i, &x, 8); }
    if (i < n) { uint64_t x = rng_next(&r); memcpy(out + i, &x, n - i); }
    return 0;
}

int kbwt_gen_csvcols(uint64_t seed, uint8_t *out, size_t n) {
    rng_t r; rng_seed(&r, seed);
    int64_t cols[8];
    for (int c = 0; c < 8; c++) cols[c] = (int64_t)rng_below(&r, 2000000) - 1000000;
    char line[256];
    size_t total = 0;
    while (total < n) {
        int len = csv_line(&r, cols, line);
        size_t k = (size_t)len < n - total ? (size_t)len : n - total;
        memcpy(out + total, line, k);
        total += k;
    }
    return 0;
}

int kbwt_gen_logs(uint64_t seed, uint8_t *out, size_t n) {
    static const char *lvl[] = {"INFO", "INFO", "INFO", "WARN", "DEBUG", "ERROR"};
    static const char *verb[] = {"GET", "GET", "POST", "PUT", "DELETE"};
    static const char *path[] = {"/api/v1/items", "/api/v1/users", "/static/app.js", "/healthz", "/api/v2/orders"};
    rng_t r; rng_seed(&r, seed);
    uint64_t ts = 1700000000000ull + rng_below(&r, 100000000000ull);
    char line[256];
    size_t total = 0;
    while (total < n) {
        ts += 1 + rng_below(&r, 50);
        int code = rng_below(&r, 20) ? 200 : (rng_below(&r, 2) ? 404 : 500);
        int len = snprintf(line, sizeof line, "%llu %s [worker-%d] %u.%u.%u.%u %s %s/%u status=%d latency_ms=%u\n",
                           (unsigned long long)ts, lvl[rng_below(&r, 6)], (int)rng_below(&r, 32), 10u,
                           (unsigned)rng_below(&r, 4), (unsigned)rng_below(&r, 256), (unsigned)rng_below(&r, 256),
                           verb[rng_below(&r, 5)], path[rng_below(&r, 5)], (unsigned)rng_below(&r, 100000), code,
                           (unsigned)(1 + rng_below(&r, 400)));
        size_t k = (size_t)len < n - total ? (size_t)len : n - total;
        memcpy(out + total, line, k);
        total += k;
    }
    return 0;
}

int kbwt_gen_bf16(uint64_t seed, uint8_t *out, size_t n) {
    rng_t r; rng_seed(&r, seed);
    for (size_t i = 0; i < n; i += 2) {
        float f = (float)(rng_normal(&r) * 0.02);
        uint32_t u; memcpy(&u, &f, 4);
        uint32_t h = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
        out[i] = (uint8_t)h;
        if (i + 1 < n) out[i + 1] = (uint8_t)(h >> 8);
    }
    return 0;
}

/* ------------------------------------------------------------------ kolmo's test inputs (tests/test_bwt.c:24-45) */

int kbwt_gen_bsc_kind(int kind, uint8_t *in, int n) {
    static const char *words[] = {"the ", "block ", "sorting ", "kolmo ", "data ", "\n", "run ", "of ", "a ", "1234 "};
    uint64_t x = 88172645463325252ull + (uint64_t)kind;
    int w = 0, wp = 0;
    for (int i = 0; i < n; i++) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        switch (kind) {
        case 0: in[i] = 0; break;
        case 1: in[i] = (uint8_t)x; break;
        case 2: in[i] = (x & 3) ? 0xf2 : (uint8_t)(x >> 8); break;
        case 3: in[i] = (uint8_t)("abcabcabd"[i % 9]); break;
        case 4: in[i] = (uint8_t)(x % 3); break;
        case 5: in[i] = (uint8_t)((i / 1000) & 0xff); break;
        case 6: in[i] = i < n / 2 ? (uint8_t)x : in[i - n / 2]; break;
        case 7: in[i] = (i % 97 == 0) ? 0xf2 : (uint8_t)("the quick brown fox "[i % 20]); break;
        default:
            if (!words[w][wp]) {
                w = (int)(x % 10); wp = 0;
                if ((x >> 20) % 50 == 0 && i > 5000) {
                    int L = 200 + (int)(x >> 40) % 300, s = i - 4000;
                    for (int k = 0; k < L && i < n; k++) in[i++] = in[s + k];
                    i--;
                    continue;
                }
            }
            in[i] = (uint8_t)words[w][wp++];
            break;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ dispatch and batches */

/* kinds: 0 markov2, 1 mixed, 2 repeats, 3 dna, 4 random, 5 csvcols, 6 logs, 7 bf16, 8 mixed text lines only.
   model_seed draws the Markov rows (blocks of one batch share their source statistics), seed everything else. */
int kbwt_gen_kind2(int kind, uint64_t model_seed, uint64_t seed, uint8_t *out, size_t n) {
    switch (kind) {
    case 0: return kbwt_gen_markov2(model_seed, seed, out, n);
    case 1: return kbwt_gen_mixed2(model_seed, seed, out, n);
    case 2: return kbwt_gen_repeats(seed, out, n);
    case 3: return kbwt_gen_dna(seed, out, n);
    case 4: return kbwt_gen_random(seed, out, n);
    case 5: return kbwt_gen_csvcols(seed, out, n);
    case 6: return kbwt_gen_logs(seed, out, n);
    case 7: return kbwt_gen_bf16(seed, out, n);
    case 8: {
        const markov3_t *m = m3_get(model_seed);
        if (!m) return -2;
        rng_t r; rng_seed(&r, seed);
        int ctx = 0;
        uint8_t line[256];
        size_t total = 0;
        while (total < n) {
            int len = m3_line(m, &r, &ctx, line);
            size_t k = (size_t)len < n - total ? (size_t)len : n - total;
            memcpy(out + total, line, k);
            total += k;
        }
        return 0;
    }
    default: return -1;
    }
}
int kbwt_gen_kind(int kind, uint64_t seed, uint8_t *out, size_t n) { return kbwt_gen_kind2(kind, seed, seed, out, n); }

/* count block sizes log-uniform in [lo, hi], each with a kind drawn uniformly from kinds[0..nkinds) and a seed */
int kbwt_gen_batch_plan(uint64_t seed, int count, int64_t lo, int64_t hi, const int *kinds, int nkinds, int64_t *sizes,
                        int *kind_out, uint64_t *seeds) {
    if (count < 0 || lo < 1 || hi < lo || nkinds < 1) return -1;
    rng_t r; rng_seed(&r, seed);
    const double a = log((double)lo), b = log((double)hi);
    for (int i = 0; i < count; i++) {
        int64_t s = (int64_t)exp(a + (b - a) * rng_unif(&r));
        if (s < lo) s = lo;
        if (s > hi) s = hi;
        sizes[i] = s;
        kind_out[i] = kinds[rng_below(&r, (uint64_t)nkinds)];
        seeds[i] = rng_next(&r);
    }
    return 0;
}
