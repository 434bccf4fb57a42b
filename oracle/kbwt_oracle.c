/* CPU oracle for the block-sorting path: test infrastructure, never the product.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this file's library, and there only
 * as the checker. The product (kolmo_amd/lib/libkbwt_hip.so) never links or calls it.
 *
 * What it restates (citations are into the kolmo 0.0.2 tree):
 *   - oracle_sa          : suffix array with the implicit end marker (shorter suffix first), by SA-IS, the algorithm
 *                          libsais runs in libsais_main / libsais_main_8u (src/algo/sais/sais.c:8747-8806). Written
 *                          from the published algorithm (Nong, Zhang, Chan 2009), not from libsais' code.
 *   - oracle_bwt_aux     : kolmo_sais_bwt_aux (sais.c:12027-12053): validation, the n <= 1 cases, U[0] = T[n-1], the
 *                          primary row skipped (libsais_bwt_copy_8u around I[0], sais.c:12043-12049), and the sampled
 *                          indexes I[k] = 1 + rank of suffix k*r that final_bwt_aux_scan_*_8u record
 *                          (sais.c:6971-7015, 7467).
 *   - oracle_unbwt_aux   : kolmo_sais_unbwt_aux (sais.c:12055-12077): validation, then libsais_unbwt_main's decode
 *                          (sais.c:11949-12022): a psi table built by counting (calculate_biPSI, sais.c:9391, restated
 *                          for single symbols) and one chain per sampled index, chain t decoding text positions
 *                          [t*r, (t+1)*r) forward from row I[t] (unbwt_decode, sais.c:10679).
 *   - oracle_unbwt       : kolmo_sais_unbwt (sais.c:12079-12082): r = n.
 *   - oracle_bwt_mod / oracle_codec18_rlog : the sampling-interval rules of src/codecs/bwt_stream.c:40-44 (codec 5)
 *                          and :224-228 (codec 18).
 *
 * How each function is pinned (tests/test_oracle.py): oracle_sa against brute-force suffix sorting of every string over
 * {a,b} up to length 12 and random strings up to 2 KiB; oracle_bwt_aux and both inverses against the reference's own
 * sais.c compiled into oracle/_ref (byte equality on the edge plan of tests/test_gpu_sa.c:74-98 and random inputs)
 * and against committed SHA-256 fixtures written from that build; the whole codec-5 stream through the libbsc 3.4.0
 * golden hashes of tests/test_bwt_golden.h; the r rules against their closed forms. No function is left unpinned. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ SA-IS with an implicit end marker */

/* bucket heads (end = 0) or one-past-ends (end = 1) of alphabet [0, K) */
static void buckets(const int32_t *s, int32_t n, int32_t K, int32_t *bkt, int end) {
    memset(bkt, 0, (size_t)K * sizeof(int32_t));
    for (int32_t i = 0; i < n; i++) bkt[s[i]]++;
    int32_t sum = 0;
    for (int32_t c = 0; c < K; c++) { sum += bkt[c]; bkt[c] = end ? sum : sum - bkt[c]; }
}

/* induced sorting: L-type suffixes left to right from the end marker, then S-type suffixes right to left.
   type[i] = 1 for S-type. The marker (suffix n) is the smallest suffix and precedes suffix n-1, which is L-type. */
static void induce(const int32_t *s, int32_t *SA, int32_t n, int32_t K, const uint8_t *type, int32_t *bkt) {
    buckets(s, n, K, bkt, 0);
    SA[bkt[s[n - 1]]++] = n - 1;
    for (int32_t i = 0; i < n; i++) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && !type[j]) SA[bkt[s[j]]++] = j;
    }
    buckets(s, n, K, bkt, 1);
    for (int32_t i = n - 1; i >= 0; i--) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && type[j]) SA[--bkt[s[j]]] = j;
    }
}

#define IS_LMS(i) ((i) > 0 && type[i] && !type[(i) - 1])

static int sais_rec(const int32_t *s, int32_t *SA, int32_t n, int32_t K) {
    if (n == 0) return 0;
    if (n == 1) { SA[0] = 0; return 0; }
    uint8_t *type = (uint8_t *)malloc((size_t)n);
    int32_t *bkt = (int32_t *)malloc((size_t)K * sizeof(int32_t));
    if (!type || !bkt) { free(type); free(bkt); return -2; }
    type[n - 1] = 0;
    for (int32_t i = n - 2; i >= 0; i--) type[i] = (s[i] < s[i + 1] || (s[i] == s[i + 1] && type[i + 1])) ? 1 : 0;

    /* stage 1: sort the LMS substrings */
    buckets(s, n, K, bkt, 1);
    for (int32_t i = 0; i < n; i++) SA[i] = -1;
    for (int32_t i = 1; i < n; i++) if (IS_LMS(i)) SA[--bkt[s[i]]] = i;
    induce(s, SA, n, K, type, bkt);

    int32_t n1 = 0;
    for (int32_t i = 0; i < n; i++) if (IS_LMS(SA[i])) SA[n1++] = SA[i];
    for (int32_t i = n1; i < n; i++) SA[i] = -1;
    /* name the LMS substrings; one that reaches the end marker is unique */
    int32_t name = 0, prev = -1;
    for (int32_t i = 0; i < n1; i++) {
        int32_t pos = SA[i], diff = 0;
        for (int32_t d = 0;; d++) {
            if (prev < 0 || pos + d == n || prev + d == n || s[pos + d] != s[prev + d] || type[pos + d] != type[prev + d]) {
                diff = 1; break;
            }
            if (d > 0 && (IS_LMS(pos + d) || IS_LMS(prev + d))) break;
        }
        if (diff) { name++; prev = pos; }
        SA[n1 + pos / 2] = name - 1;
    }
    for (int32_t i = n - 1, j = n - 1; i >= n1; i--) if (SA[i] >= 0) SA[j--] = SA[i];

    /* stage 2: the order of the LMS suffixes, by recursion when the names repeat */
    int32_t *s1 = SA + n - n1;
    int rc = 0;
    if (name < n1) rc = sais_rec(s1, SA, n1, name);
    else for (int32_t i = 0; i < n1; i++) SA[s1[i]] = i;
    if (rc) { free(type); free(bkt); return rc; }

    /* stage 3: place the sorted LMS suffixes at their bucket ends and induce the rest */
    for (int32_t i = 1, j = 0; i < n; i++) if (IS_LMS(i)) s1[j++] = i;
    for (int32_t i = 0; i < n1; i++) SA[i] = s1[SA[i]];
    for (int32_t i = n1; i < n; i++) SA[i] = -1;
    buckets(s, n, K, bkt, 1);
    for (int32_t i = n1 - 1; i >= 0; i--) { int32_t j = SA[i]; SA[i] = -1; SA[--bkt[s[j]]] = j; }
    induce(s, SA, n, K, type, bkt);
    free(type); free(bkt);
    return 0;
}
#undef IS_LMS

/* Suffix array of T[0..n) (libsais(T, SA, n, 0, NULL)): 0, or -2 when out of memory. */
int32_t oracle_sa(const uint8_t *T, int32_t *SA, int32_t n) {
    if (n <= 0) return 0;
    int32_t *s = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!s) return -2;
    for (int32_t i = 0; i < n; i++) s[i] = T[i];
    int rc = sais_rec(s, SA, n, 256);
    free(s);
    return rc;
}

/* ------------------------------------------------------------------ forward BWT with sampled indexes */

int32_t oracle_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I) {
    if (T == NULL || U == NULL || n < 0 || r < 2 || (r & (r - 1)) != 0 || I == NULL) return -1;
    if (n <= 1) { if (n == 1) U[0] = T[0]; I[0] = n; return 0; }
    int32_t *SA = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    uint8_t *out = (uint8_t *)malloc((size_t)n);
    if (!SA || !out || oracle_sa(T, SA, n)) { free(SA); free(out); return -2; }
    int32_t p0 = -1;
    for (int32_t p = 0; p < n; p++) {
        int32_t i = SA[p];
        if ((i & (r - 1)) == 0) I[i / r] = p + 1;
        if (i == 0) p0 = p;
    }
    out[0] = T[n - 1];
    for (int32_t p = 0; p < n; p++) {
        if (p == p0) continue;
        out[p < p0 ? p + 1 : p] = T[SA[p] - 1];
    }
    memcpy(U, out, (size_t)n);
    free(SA); free(out);
    return 0;
}

/* ------------------------------------------------------------------ inverse BWT */

int32_t oracle_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, const int32_t *I) {
    if (T == NULL || U == NULL || n < 0 || (r != n && (r < 2 || (r & (r - 1)) != 0)) || I == NULL) return -1;
    if (n <= 1) { if (I[0] != n) return -1; if (n == 1) U[0] = T[0]; return 0; }
    for (int64_t t = 0; t <= (n - 1) / r; t++) if (I[t] <= 0 || I[t] > n) return -1;
    int64_t count[256] = {0}, C[257];
    for (int32_t i = 0; i < n; i++) count[T[i]]++;
    C[0] = 1;                                               /* row 0 of the first column is the end marker */
    for (int c = 0; c < 256; c++) C[c + 1] = C[c] + count[c];
    uint32_t *psi = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    uint8_t *out = (uint8_t *)malloc((size_t)n);
    if (!psi || !out) { free(psi); free(out); return -2; }
    /* rows of the full BWT: the marker sits at row I[0]; LF(j) = C[c] + occ(c, j); psi is its inverse */
    int64_t next[256];
    for (int c = 0; c < 256; c++) next[c] = C[c];
    const int32_t p0 = I[0];
    for (int64_t j = 0; j <= n; j++) {
        if (j == p0) continue;
        uint8_t c = T[j < p0 ? j : j - 1];
        psi[next[c]++] = (uint32_t)j;
    }
    psi[0] = (uint32_t)p0;
    /* first-column byte of a row: the c with C[c] <= row < C[c+1] */
    for (int64_t t = 0; t <= (n - 1) / r; t++) {
        int64_t row = I[t], end = (t + 1) * (int64_t)r < n ? (t + 1) * (int64_t)r : n;
        for (int64_t p = t * (int64_t)r; p < end; p++) {
            int lo = 0, hi = 255;
            while (lo < hi) { int mid = (lo + hi + 1) / 2; if (C[mid] <= row) lo = mid; else hi = mid - 1; }
            out[p] = (uint8_t)lo;
            row = psi[row];
        }
    }
    memcpy(U, out, (size_t)n);
    free(psi); free(out);
    return 0;
}

int32_t oracle_unbwt(const uint8_t *T, uint8_t *U, int32_t n, int32_t i) { return oracle_unbwt_aux(T, U, n, n, &i); }

/* ------------------------------------------------------------------ sampling-interval rules */

/* codec 5 (bwt_stream.c:40-44): the largest power of two <= n/8, minus one */
int32_t oracle_bwt_mod(int32_t n) {
    int32_t x = n / 8, p = 1;
    if (x <= 0) return 0;
    while ((int64_t)p * 2 <= x) p *= 2;
    return p - 1;
}

/* codec 18 (bwt_stream.c:224-228): log2 of r with index_log2 = 0; at most 255 sampled indexes */
int32_t oracle_codec18_rlog(int32_t n) {
    int32_t r = oracle_bwt_mod(n) + 1, rlog = 0;
    while ((1 << (rlog + 1)) <= r) rlog++;
    if (rlog < 1) rlog = 1;
    while (rlog < 30 && (((int64_t)n - 1) >> rlog) > 255) rlog++;
    return rlog;
}
