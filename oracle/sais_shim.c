/* Test infrastructure: kolmo's three kolmo_sais_* entry points (src/algo/sais/sais.h:13,17,20) answered by the CPU
 * oracle, so that the reference's own code (the codec glue bwt_stream.c, lzp.c, qlfc.c, or the whole library),
 * linked against this file instead of sais.c, runs its streams over the oracle. Used only by tests/ to pin the oracle
 * through the libbsc golden hashes and kolmo's determinism digests. oracle_shim_calls counts the calls, so a test
 * can show that the block-sorting path was taken. */
#include <stdint.h>
#include <stddef.h>

int32_t oracle_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I);
int32_t oracle_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, const int32_t *I);

static int64_t g_calls[2];   /* forward, inverse */

int64_t oracle_shim_calls(int32_t which) { return __atomic_load_n(&g_calls[which & 1], __ATOMIC_RELAXED); }

int32_t kolmo_sais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, int32_t *I, int32_t threads) {
    (void)threads;
    if (A == NULL) return -1;
    __atomic_fetch_add(&g_calls[0], 1, __ATOMIC_RELAXED);
    return oracle_bwt_aux(T, U, n, r, I);
}
int32_t kolmo_sais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, const int32_t *I,
                             int32_t threads) {
    (void)threads;
    if (A == NULL) return -1;
    __atomic_fetch_add(&g_calls[1], 1, __ATOMIC_RELAXED);
    return oracle_unbwt_aux(T, U, n, r, I);
}
int32_t kolmo_sais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t i, int32_t threads) {
    return kolmo_sais_unbwt_aux(T, U, A, n, n, &i, threads);
}
