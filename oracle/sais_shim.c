/* Test infrastructure: kolmo's three kolmo_sais_* entry points (src/algo/sais/sais.h:13,17,20) answered by the CPU
 * oracle, so that the reference's own codec glue (bwt_stream.c, lzp.c, qlfc.c), linked against this file instead of
 * sais.c, runs its streams over the oracle. Used only by tests/ to pin the oracle through the libbsc golden hashes. */
#include <stdint.h>
#include <stddef.h>

int32_t oracle_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I);
int32_t oracle_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, const int32_t *I);

int32_t kolmo_sais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, int32_t *I, int32_t threads) {
    (void)threads;
    if (A == NULL) return -1;
    return oracle_bwt_aux(T, U, n, r, I);
}
int32_t kolmo_sais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, const int32_t *I,
                             int32_t threads) {
    (void)threads;
    if (A == NULL) return -1;
    return oracle_unbwt_aux(T, U, n, r, I);
}
int32_t kolmo_sais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t i, int32_t threads) {
    return kolmo_sais_unbwt_aux(T, U, A, n, n, &i, threads);
}
