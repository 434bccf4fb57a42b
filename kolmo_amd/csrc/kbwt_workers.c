// kolmo's block workers in C, to measure concurrent calls of the drop-in without an interpreter in the way.
//
// `workers` threads take the blocks of one buffer in turn from a shared counter, as kolmo's workers take jobs
// (src/core/compress.c:231-256, src/core/decode.c:404-410), and run kolmo_sais_bwt_aux and then
// kolmo_sais_unbwt_aux on each block from host buffers. Every round trip is compared with its block.
#include <pthread.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kbwt.h"

#define KBWT_WORKERS_API __attribute__((visibility("default")))

// The first failure of the last run: which call, its return code and the library's message.
static char g_fail[256];
static int32_t g_fail_set;
static void note_failure(const char *what, int32_t block, int32_t rc) {
    if (__atomic_exchange_n(&g_fail_set, 1, __ATOMIC_ACQ_REL)) return;
    snprintf(g_fail, sizeof g_fail, "%s of block %d: rc %d %s", what, (int)block, (int)rc, rc ? kbwt_last_error() : "");
}

// The first failure of the last kbwt_workers_run that returned -2, or "".
KBWT_WORKERS_API const char *kbwt_workers_last_error(void) { return g_fail; }

typedef struct {
    const uint8_t *T;
    int64_t n;
    int32_t block;
    const int32_t *rs;   // r of every block
    int32_t nblocks;
    int32_t next;        // the next block to take
    int32_t failed;
} Job;

static void *worker(void *arg) {
    Job *j = (Job *)arg;
    const int32_t cap = j->block;
    uint8_t *U = (uint8_t *)malloc((size_t)cap), *back = (uint8_t *)malloc((size_t)cap);
    int32_t *I = (int32_t *)malloc(((size_t)cap + 1) * sizeof(int32_t));
    int32_t *A = (int32_t *)malloc((size_t)cap * sizeof(int32_t));   // the callers' scratch array (must be set)
    if (!U || !back || !I || !A) {
        __atomic_store_n(&j->failed, 1, __ATOMIC_RELAXED);
    } else {
        for (;;) {
            const int32_t k = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
            if (k >= j->nblocks) break;
            const int64_t o = (int64_t)k * j->block;
            const int32_t len = (int32_t)(j->n - o < j->block ? j->n - o : j->block);
            const uint8_t *T = j->T + o;
            int32_t rc = kolmo_sais_bwt_aux(T, U, A, len, j->rs[k], I, 1);
            if (rc != 0) note_failure("kolmo_sais_bwt_aux", k, rc);
            if (rc == 0) {
                rc = kolmo_sais_unbwt_aux(U, back, A, len, j->rs[k], I, 1);
                if (rc != 0) note_failure("kolmo_sais_unbwt_aux", k, rc);
                else if (memcmp(back, T, (size_t)len) != 0) {
                    note_failure("round trip", k, 0);
                    rc = 1;
                }
            }
            if (rc != 0) {
                __atomic_store_n(&j->failed, 1, __ATOMIC_RELAXED);
                break;
            }
        }
    }
    free(U);
    free(back);
    free(I);
    free(A);
    return NULL;
}

// Best wall time in seconds over `reps` passes of `workers` threads over the n bytes of T cut into blocks of `block`
// bytes (the last one shorter; rs[k] is block k's sampling interval). -1 on bad arguments, -2 when a call fails or a
// round trip differs from its block.
KBWT_WORKERS_API double kbwt_workers_run(const uint8_t *T, int64_t n, int32_t block, const int32_t *rs,
                                         int32_t workers, int32_t reps) {
    if (!T || !rs || n < 2 || block < 2 || workers < 1 || workers > 256 || reps < 1) return -1.0;
    const int64_t nb = (n + block - 1) / block;
    if (nb > (1 << 20) || n - (nb - 1) * block < 2) return -1.0;
    pthread_t th[256];
    double best = -1.0;
    g_fail_set = 0;
    g_fail[0] = 0;
    for (int32_t rep = 0; rep < reps; rep++) {
        Job j = {T, n, block, rs, (int32_t)nb, 0, 0};
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        int32_t started = 0;
        for (; started < workers; started++)
            if (pthread_create(&th[started], NULL, worker, &j) != 0) break;
        if (started == 0) return -1.0;
        for (int32_t w = 0; w < started; w++) pthread_join(th[w], NULL);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        if (j.failed || j.next < j.nblocks) return -2.0;
        const double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        if (best < 0 || s < best) best = s;
    }
    return best;
}
