// Kernels of the forward suffix sort (prefix doubling), the BWT emission and the inverse BWT, for gfx950.
//
// Forward, for a block T[0..n):
//   1. k_byte_hist + k_reduce_hist: byte histogram; the host turns it into a dense order-preserving symbol code of
//      w = ceil(log2 sigma) bits and picks h0 symbols per first key (h0*w <= 64 bits).
//   2. k_hist_T + k_scan_digits: the 8-bit digit histograms of every first key, in one read of T; exclusive scans
//      give each digit's global bucket offsets (onesweep).
//   3. k_sort_pass: one stable LSD pass per varying digit. The first pass builds keys from T directly. A tile of
//      BLOCK*ITEMS elements is ranked by wave64 ballot multisplit, its digit offsets come from a decoupled lookback
//      over earlier tiles, and it is written through LDS so equal digits leave as contiguous runs.
//   4. k_regroup: group heads (rank = SA position of the group's first row) by a max-scan, ISA[suffix] = head,
//      and compaction of the rows still in groups of two or more, all with decoupled lookback.
//   5. per doubling round with offset h: k_round_keys gathers (group head, ISA[i+h] + h or n-1-i past the end) for
//      the active rows, the same radix passes sort them, k_regroup splits the groups.
//   6. k_bwt + k_samples: U in libsais' convention and I[k] = 1 + ISA[k*r].
// Inverse, for a BWT L[0..n) with primary index P (the end marker's row in the full n+1-row matrix):
//   1. byte histogram -> C[c] = 1 + #bytes < c.
//   2. k_lf_pass: LF(row) = C[c] + occ(c, row) for every row, a stable multisplit ranking with lookback.
//   3. k_walk_count: from every splitter row (row % D == 0, and P) follow LF to the next splitter, recording the
//      sublist length; k_rank_jump: synchronous pointer jumping gives each splitter's text position;
//      k_walk_emit: each sublist is walked again and its bytes written.
#include "kbwt_device.h"
#include "kbwt_internal.h"

namespace kbwt {

// ------------------------------------------------------------------ histograms
//
// Every histogram kernel accumulates in LDS and adds its non-zero bins to a global `counts` array with one atomic
// per bin; k_scan_digits turns counts into exclusive scans and clears them again, so counts is all zero between
// uses (it is zeroed once when allocated).

__global__ __launch_bounds__(256) void k_byte_hist(const uint8_t *__restrict__ T, uint32_t n,
                                                   uint32_t *__restrict__ counts) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 1024; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const size_t n4 = n / 4;
    const uint32_t *T4 = reinterpret_cast<const uint32_t *>(T);
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n4; e += stride) {
        uint32_t x = T4[e];
        atomicAdd(&h[wave][x & 255u], 1u);
        atomicAdd(&h[wave][(x >> 8) & 255u], 1u);
        atomicAdd(&h[wave][(x >> 16) & 255u], 1u);
        atomicAdd(&h[wave][x >> 24], 1u);
    }
    if (blockIdx.x == 0)
        for (size_t e = n4 * 4 + threadIdx.x; e < n; e += 256) atomicAdd(&h[wave][T[e]], 1u);
    __syncthreads();
    uint32_t c = h[0][threadIdx.x] + h[1][threadIdx.x] + h[2][threadIdx.x] + h[3][threadIdx.x];
    if (c) atomicAdd(&counts[threadIdx.x], c);
}

// One block per histogram set: tot = counts, excl = base + exclusive scan of counts; counts is cleared.
__global__ __launch_bounds__(256) void k_scan_digits(uint32_t *__restrict__ counts, uint32_t *__restrict__ tot,
                                                     uint32_t *__restrict__ excl, uint32_t base) {
    __shared__ uint32_t scratch[8];
    const uint32_t set = blockIdx.x, b = threadIdx.x;
    const uint32_t s = counts[set * 256 + b];
    uint32_t total;
    uint32_t ex = block_excl_sum<256>(s, scratch, &total);
    if (tot) tot[set * 256 + b] = s;
    if (excl) excl[set * 256 + b] = base + ex;
    counts[set * 256 + b] = 0;
}

// First key of suffix i: h0 symbol codes of w bits (zero past the end) above the 8-bit BWT byte T[i-1] (0 for i = 0,
// the primary row, whose byte is never emitted). c points at the code of T[i], raw at T[i-1].
__device__ __forceinline__ uint64_t build_key(const uint8_t *c, uint8_t bwt_byte, uint32_t h0, uint32_t w) {
    uint64_t k = 0;
    for (uint32_t b = 0; b < h0; b++) k = (k << w) | c[b];
    return (k << 8) | bwt_byte;
}

// Digit histograms of the first keys, P digits of 8 bits above the BWT byte.
__global__ __launch_bounds__(256) void k_hist_T(const uint8_t *__restrict__ T, uint32_t n,
                                                const uint8_t *__restrict__ code, uint32_t w, uint32_t h0,
                                                uint32_t P, uint32_t *__restrict__ counts) {
    constexpr int HT = 4096;
    __shared__ uint32_t h[MAXP][256];
    __shared__ uint8_t s_code[256];
    __shared__ uint8_t s_c[HT + 64];
    for (int i = threadIdx.x; i < MAXP * 256; i += 256) (&h[0][0])[i] = 0;
    s_code[threadIdx.x] = code[threadIdx.x];
    const uint32_t ntiles = (n + HT - 1) / HT;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const size_t base = (size_t)tile * HT;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < HT + h0; j += 256) {
            size_t e = base + j;
            s_c[j] = e < n ? s_code[T[e]] : 0;
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < HT; j += 256) {
            if (base + j >= n) break;
            uint64_t key = build_key(s_c + j, 0, h0, w);
            for (uint32_t p = 0; p < P; p++) atomicAdd(&h[p][(uint32_t)(key >> (8 * p + 8)) & 255u], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < P * 256; i += 256) {
        uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&counts[i], c);
    }
}

// Digit histograms of stored keys at the given shifts.
__global__ __launch_bounds__(256) void k_hist_keys(const uint64_t *__restrict__ K, uint32_t m, Shifts sh,
                                                   uint32_t *__restrict__ counts) {
    __shared__ uint32_t h[MAXP][256];
    for (int i = threadIdx.x; i < MAXP * 256; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < m; e += stride) {
        uint64_t key = K[e];
        for (uint32_t p = 0; p < sh.P; p++) atomicAdd(&h[p][(uint32_t)(key >> sh.s[p]) & 255u], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < sh.P * 256; i += 256) {
        uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&counts[i], c);
    }
}

// ------------------------------------------------------------------ onesweep radix pass

// Stable pass over m (key, value) pairs on the 8-bit digit at `shift`. FROM_T: keys are built from T (value = the
// suffix index). dbase[d]: global start of digit d. status: ntiles*256 lookback words; counter: tile-id counter.
// The tile leaves through one LDS buffer, keys first and then values, so equal digits go out as contiguous runs.
template <int BLOCK, int ITEMS, bool FROM_T>
__global__ __launch_bounds__(BLOCK) void k_sort_pass(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                     uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     uint32_t m, uint32_t shift, const uint32_t *__restrict__ dbase,
                                                     uint64_t *__restrict__ status, uint32_t *__restrict__ counter,
                                                     uint32_t epoch, uint32_t ntiles, const uint8_t *__restrict__ T,
                                                     const uint8_t *__restrict__ code, uint32_t w, uint32_t h0) {
    constexpr int W = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(BLOCK >= 256, "one thread per digit");
    static_assert(TILE * 8 >= 2 * TILE + 128, "LDS staging of raw bytes and codes");
    __shared__ uint32_t s_wcnt[W][256];
    __shared__ uint32_t s_tstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_scratch[16];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_buf[TILE];
    __shared__ uint8_t s_code[FROM_T ? 256 : 1];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    for (int i = tid; i < W * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    if (FROM_T && tid < 256) s_code[tid] = code[tid];
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * TILE;

    uint64_t key[ITEMS];
    uint32_t val[ITEMS];
    if (FROM_T) {
        // raw[j] = T[base + j - 1] (the BWT byte of element j), cod[j] = code of T[base + j], zero past the end
        uint8_t *s_raw = reinterpret_cast<uint8_t *>(s_buf);
        uint8_t *s_cod = s_raw + TILE + 64;
        for (uint32_t j = tid; j < (uint32_t)TILE + h0; j += BLOCK) {
            size_t e = base + j;
            uint8_t t = e < m ? T[e] : 0;
            s_cod[j] = e < m ? s_code[t] : 0;
            if (j + 1 <= (uint32_t)TILE) s_raw[j + 1] = t;
        }
        if (tid == 0) s_raw[0] = base > 0 ? T[base - 1] : 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint32_t local = wave * WAVE * ITEMS + k * WAVE + lane;
            key[k] = build_key(s_cod + local, s_raw[local], h0, w);
            val[k] = (uint32_t)(base + local);
        }
        __syncthreads();   // s_buf is reused below
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
            if (e < m) { key[k] = kin[e]; val[k] = vin[e]; } else { key[k] = 0; val[k] = 0; }
        }
    }

    // stable ranking inside each wave, items in element order (k-major, lane-minor)
    uint32_t rk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        bool valid = e < m;
        uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
        uint64_t peers = match_digit8(d, __ballot(valid));
        rk[k] = 0;
        if (valid) {
            uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
            uint32_t c = s_wcnt[wave][d];
            rk[k] = c + before;
            s_wcnt[wave][d] = c + (uint32_t)__popcll(peers);   // every peer writes the same value
        }
    }
    __syncthreads();

    // per digit: wave prefixes, tile total, lookback over earlier tiles
    uint32_t tot = 0;
    if (tid < 256) {
#pragma unroll
        for (int w2 = 0; w2 < W; w2++) { uint32_t c = s_wcnt[w2][tid]; s_wcnt[w2][tid] = tot; tot += c; }
        if (tile == 0) st_store(status + tid, st_pack(epoch, ST_INC, tot));
        else st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_AGG, tot));
    }
    uint32_t tstart = block_excl_sum<BLOCK>(tid < 256 ? tot : 0u, s_scratch, nullptr);
    if (tid < 256) {
        uint32_t excl = 0;
        if (tile > 0) {
            excl = lookback<false>(status + tid, 256, tile, epoch);
            st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_INC, excl + tot));
        }
        s_tstart[tid] = tstart;
        s_gbase[tid] = dbase[tid] + excl - tstart;
    }
    __syncthreads();

    // local sorted position of each item
    uint32_t lidx[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
        lidx[k] = s_tstart[d] + s_wcnt[wave][d] + rk[k];
    }
    const uint32_t cnt = (uint32_t)((m - base) < (size_t)TILE ? (m - base) : (size_t)TILE);
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        if (e < m) s_buf[lidx[k]] = key[k];
    }
    __syncthreads();
    uint32_t dg[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t j = tid + k * BLOCK;
        dg[k] = 0;
        if (j < cnt) {
            uint64_t kk = s_buf[j];
            dg[k] = (uint32_t)(kk >> shift) & 255u;
            kout[s_gbase[dg[k]] + j] = kk;
        }
    }
    __syncthreads();
    uint32_t *s_v = reinterpret_cast<uint32_t *>(s_buf);
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        if (e < m) s_v[lidx[k]] = val[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t j = tid + k * BLOCK;
        if (j < cnt) vout[s_gbase[dg[k]] + j] = s_v[j];
    }
}

// ------------------------------------------------------------------ regroup (group heads, ISA, compaction)

// Sorted keys K[0..m) with values V (suffixes). Row a goes to SA position pos[a] (pos == nullptr: a). Rows a-1 and a
// share a group when their keys agree above bit kshift. Writes SA[pos[a]] = V[a] (when SA != nullptr), ISA[V[a]] =
// head of its group (SA position of the group's first row), the BWT byte of row pos[a] into L (the key's low byte
// when T == nullptr, else T[V[a]-1]), and compacts the rows of groups of two or more into (pos_out, grp_out); the
// last tile writes their count to *m_out.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_regroup(const uint64_t *__restrict__ K, const uint32_t *__restrict__ V,
                                                   const uint32_t *__restrict__ pos, uint32_t m, uint32_t kshift,
                                                   uint32_t *__restrict__ SA, uint32_t *__restrict__ ISA,
                                                   uint8_t *__restrict__ L, const uint8_t *__restrict__ T,
                                                   uint32_t *__restrict__ pos_out, uint32_t *__restrict__ grp_out,
                                                   uint32_t *__restrict__ m_out, uint64_t *__restrict__ st_sum,
                                                   uint64_t *__restrict__ st_max, uint32_t *__restrict__ counter,
                                                   uint32_t epoch, uint32_t ntiles) {
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ uint32_t s_scratch[32];
    __shared__ uint32_t s_tile, s_ex_sum, s_ex_max;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * TILE;

    uint32_t lkeep[ITEMS], lhead[ITEMS], lp[ITEMS];
    bool lkept[ITEMS];
    uint8_t lbyte[ITEMS];
    uint32_t run_sum = 0, run_max = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t a = base + (size_t)k * BLOCK + tid;
        bool valid = a < m;
        bool flag = false, keep = false;
        uint32_t p = 0;
        lbyte[k] = 0;
        if (valid) {
            uint64_t kr = K[a];
            uint64_t kc = kr >> kshift;
            flag = (a == 0) || ((K[a - 1] >> kshift) != kc);
            bool flagn = (a + 1 >= m) || ((K[a + 1] >> kshift) != kc);
            keep = !(flag && flagn);
            p = pos ? pos[a] : (uint32_t)a;
            lbyte[k] = (uint8_t)kr;
        }
        uint32_t ex_s, in_m, tot_s, tot_m;
        block_scan_sum_max<BLOCK>(keep ? 1u : 0u, flag ? p : 0u, ex_s, in_m, tot_s, tot_m, s_scratch);
        lkeep[k] = run_sum + ex_s;
        lhead[k] = run_max > in_m ? run_max : in_m;
        lkept[k] = keep;
        lp[k] = p;
        run_sum += tot_s;
        run_max = run_max > tot_m ? run_max : tot_m;
    }
    if (tid == 0) {
        uint32_t ex_s = 0, ex_m = 0;
        if (tile == 0) {
            st_store(st_sum, st_pack(epoch, ST_INC, run_sum));
            st_store(st_max, st_pack(epoch, ST_INC, run_max));
        } else {
            st_store(st_sum + tile, st_pack(epoch, ST_AGG, run_sum));
            st_store(st_max + tile, st_pack(epoch, ST_AGG, run_max));
            ex_s = lookback<false>(st_sum, 1, tile, epoch);
            ex_m = lookback<true>(st_max, 1, tile, epoch);
            st_store(st_sum + tile, st_pack(epoch, ST_INC, ex_s + run_sum));
            uint32_t im = ex_m > run_max ? ex_m : run_max;
            st_store(st_max + tile, st_pack(epoch, ST_INC, im));
        }
        s_ex_sum = ex_s;
        s_ex_max = ex_m;
        if (tile == ntiles - 1) *m_out = ex_s + run_sum;
    }
    __syncthreads();
    const uint32_t ex_sum = s_ex_sum, ex_max = s_ex_max;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t a = base + (size_t)k * BLOCK + tid;
        if (a < m) {
            uint32_t head = ex_max > lhead[k] ? ex_max : lhead[k];
            uint32_t i = V[a];
            if (SA) SA[lp[k]] = i;
            ISA[i] = head;
            L[lp[k]] = T ? (i ? T[i - 1] : (uint8_t)0) : lbyte[k];
            if (lkept[k]) {
                uint32_t o = ex_sum + lkeep[k];
                pos_out[o] = lp[k];
                grp_out[o] = head;
            }
        }
    }
}

// ------------------------------------------------------------------ doubling round keys

// key = (group head << 32) | sec, sec = ISA[i+h] + h when i+h < n, else n-1-i (below every rank, shorter suffix
// first). orand[0] |= keys, orand[1] &= keys (which digits vary): one pair of atomics per workgroup.
__global__ __launch_bounds__(256) void k_round_keys(const uint32_t *__restrict__ pos, const uint32_t *__restrict__ grp,
                                                    uint32_t m, const uint32_t *__restrict__ SA,
                                                    const uint32_t *__restrict__ ISA, uint32_t n, uint32_t h,
                                                    uint64_t *__restrict__ K, uint32_t *__restrict__ V,
                                                    unsigned long long *__restrict__ orand) {
    __shared__ uint64_t s_o[4], s_a[4];
    uint64_t ko = 0, ka = ~0ull;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t a = (size_t)blockIdx.x * 256 + threadIdx.x; a < m; a += stride) {
        uint32_t i = SA[pos[a]];
        uint64_t j = (uint64_t)i + h;
        uint32_t sec = j < n ? ISA[j] + h : n - 1u - i;
        uint64_t key = ((uint64_t)grp[a] << 32) | sec;
        K[a] = key;
        V[a] = i;
        ko |= key;
        ka &= key;
    }
    for (int o = 1; o < WAVE; o <<= 1) {
        ko |= __shfl_xor(ko, o, WAVE);
        ka &= __shfl_xor(ka, o, WAVE);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & (WAVE - 1)) == 0) { s_o[wave] = ko; s_a[wave] = ka; }
    __syncthreads();
    if (threadIdx.x == 0) {
        ko = s_o[0] | s_o[1] | s_o[2] | s_o[3];
        ka = s_a[0] & s_a[1] & s_a[2] & s_a[3];
        atomicOr(&orand[0], (unsigned long long)ko);
        atomicAnd(&orand[1], (unsigned long long)ka);
    }
}

// ------------------------------------------------------------------ BWT emission

// U in libsais' convention from the row bytes L[p] = T[SA[p]-1]: U[0] = T[n-1], row p goes to U[p+1] before the
// primary row p0 = ISA[0] and to U[p] after it (row p0 itself, the end marker's, is dropped). Four bytes per thread.
__global__ __launch_bounds__(256) void k_bwt(const uint8_t *__restrict__ T, const uint8_t *__restrict__ L,
                                             const uint32_t *__restrict__ ISA, uint8_t *__restrict__ U, uint32_t n) {
    const uint32_t p0 = ISA[0];
    size_t o0 = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        size_t o = o0 + k;   // output byte o comes from row o-1 (o <= p0) or row o (o > p0)
        if (o >= n) break;
        U[o] = o == 0 ? T[n - 1] : (o <= p0 ? L[o - 1] : L[o]);
    }
}

__global__ void k_samples(const uint32_t *__restrict__ ISA, uint32_t r, uint32_t count, int32_t *__restrict__ I) {
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < count) I[k] = (int32_t)(ISA[k * (size_t)r] + 1u);
}

// ------------------------------------------------------------------ inverse BWT

// LF for every byte of the BWT L[0..n): lf[q] = C[L[q]] + #{q' < q : L[q'] = L[q]}. The row of L[q] in the full
// matrix is q before the primary index, q+1 from it on, so lf[q] is LF of that row. cbase[c] = C[c].
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_lf_pass(const uint8_t *__restrict__ L, uint32_t n,
                                                   const uint32_t *__restrict__ cbase, uint32_t *__restrict__ lf,
                                                   uint64_t *__restrict__ status, uint32_t *__restrict__ counter,
                                                   uint32_t epoch, uint32_t ntiles) {
    constexpr int W = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ uint32_t s_wcnt[W][256];
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_tile;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    for (int i = tid; i < W * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * TILE;
    uint32_t dg[ITEMS], rk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        bool valid = e < n;
        uint32_t d = valid ? L[e] : 0u;
        uint64_t peers = match_digit8(d, __ballot(valid));
        dg[k] = d;
        rk[k] = 0;
        if (valid) {
            uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
            uint32_t c = s_wcnt[wave][d];
            rk[k] = c + before;
            s_wcnt[wave][d] = c + (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t tot = 0;
#pragma unroll
        for (int w2 = 0; w2 < W; w2++) { uint32_t c = s_wcnt[w2][tid]; s_wcnt[w2][tid] = tot; tot += c; }
        uint32_t excl = 0;
        if (tile == 0) st_store(status + tid, st_pack(epoch, ST_INC, tot));
        else {
            st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_AGG, tot));
            excl = lookback<false>(status + tid, 256, tile, epoch);
            st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_INC, excl + tot));
        }
        s_base[tid] = cbase[tid] + excl;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        if (e < n) lf[e] = s_base[dg[k]] + s_wcnt[wave][dg[k]] + rk[k];
    }
}

// Byte of a row from its LF value: the c with C[c] <= v < C[c+1] (s_c holds C[0..255]).
__device__ __forceinline__ uint8_t row_byte(const uint32_t *s_c, uint32_t v) {
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 128; step > 0; step >>= 1)
        if (s_c[lo + step] <= v) lo += step;
    return (uint8_t)lo;
}

// Splitter s: row s*D for s < nreg, row P for s == nreg (only when P % D != 0). Walk LF from it to the next splitter
// row; link[s] = (steps << 32) | next splitter, len[s] = steps. The bytes met on the way (the text backwards from the
// splitter's suffix) go to stage[s*cap ..] for the first cap steps; cont[s] keeps the row of step cap of a longer
// sublist. The marker row P ends the list: link = (0, S).
__global__ __launch_bounds__(256) void k_walk_count(const uint32_t *__restrict__ lf, uint32_t n, uint32_t P,
                                                    uint32_t dshift, uint32_t nreg, uint32_t S,
                                                    const uint32_t *__restrict__ cbase, uint64_t *__restrict__ link,
                                                    uint32_t *__restrict__ len, uint8_t *__restrict__ stage,
                                                    uint32_t cap, uint32_t *__restrict__ cont) {
    __shared__ uint32_t s_c[256];
    s_c[threadIdx.x] = cbase[threadIdx.x];
    __syncthreads();
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s > S) return;
    if (s == S) { link[S] = S; return; }
    const uint32_t dmask = (1u << dshift) - 1u;
    const uint32_t sidP = (P & dmask) == 0 ? (P >> dshift) : nreg;
    uint32_t row = s < nreg ? (uint32_t)(s << dshift) : P;
    if (row == P) { link[s] = S; len[s] = 0; return; }
    uint32_t steps = 0;
    uint8_t *st = stage + s * (size_t)cap;
    do {
        if (steps == cap) cont[s] = row;
        uint32_t q = row < P ? row : row - 1u;
        uint32_t v = lf[q];
        if (steps < cap) st[steps] = row_byte(s_c, v);
        steps++;
        row = v;
    } while ((row & dmask) != 0 && row != P && steps <= n);
    uint32_t nxt = row == P ? sidP : (row >> dshift);
    link[s] = ((uint64_t)steps << 32) | nxt;
    len[s] = steps;
}

// One synchronous pointer-jumping step: out[s] = (acc[s] + acc[ptr], ptr[ptr]) unless ptr == S (the end).
__global__ __launch_bounds__(256) void k_rank_jump(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                   uint32_t S) {
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s > S) return;
    uint64_t v = in[s];
    uint32_t ptr = (uint32_t)v;
    if (ptr != S) {
        uint64_t w = in[ptr];
        v = (((v >> 32) + (w >> 32)) << 32) | (uint32_t)w;
    }
    out[s] = v;
}

// Write every sublist: its splitter's suffix starts at text position pos = link >> 32 (after ranking), and step t of
// the walk is the byte T[pos-1-t]. The first cap bytes come from stage; a longer sublist walks on from cont[s].
__global__ __launch_bounds__(256) void k_walk_emit(const uint32_t *__restrict__ lf, uint32_t n, uint32_t P,
                                                   uint32_t S, const uint64_t *__restrict__ link,
                                                   const uint32_t *__restrict__ len, const uint32_t *__restrict__ cbase,
                                                   const uint8_t *__restrict__ stage, uint32_t cap,
                                                   const uint32_t *__restrict__ cont, uint8_t *__restrict__ out) {
    __shared__ uint32_t s_c[256];
    s_c[threadIdx.x] = cbase[threadIdx.x];
    __syncthreads();
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= S) return;
    const uint32_t steps = len[s];
    if (steps == 0) return;
    const uint32_t pos = (uint32_t)(link[s] >> 32);
    if (pos > n || steps > pos) return;   // only for a string that is not a BWT; the caller rejects those first
    const uint8_t *st = stage + s * (size_t)cap;
    const uint32_t staged = steps < cap ? steps : cap;
    uint8_t *o = out + pos - 1;
    for (uint32_t t = 0; t < staged; t++) o[-(int64_t)t] = st[t];
    if (steps > cap) {
        uint32_t row = cont[s];
        for (uint32_t t = cap; t < steps; t++) {
            uint32_t q = row < P ? row : row - 1u;
            uint32_t v = lf[q];
            o[-(int64_t)t] = row_byte(s_c, v);
            row = v;
        }
    }
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// A plain 16-byte-vector copy: the reference point that hardware byte counters are scaled by.
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// ------------------------------------------------------------------ launch wrappers

void launch_byte_hist(const uint8_t *T, uint32_t n, uint32_t *counts, uint32_t grid, hipStream_t s) {
    k_byte_hist<<<grid, 256, 0, s>>>(T, n, counts);
}
void launch_scan_digits(uint32_t *counts, uint32_t nsets, uint32_t *tot, uint32_t *excl, uint32_t base,
                        hipStream_t s) {
    k_scan_digits<<<nsets, 256, 0, s>>>(counts, tot, excl, base);
}
void launch_hist_T(const uint8_t *T, uint32_t n, const uint8_t *code, uint32_t w, uint32_t h0, uint32_t P,
                   uint32_t *counts, uint32_t grid, hipStream_t s) {
    k_hist_T<<<grid, 256, 0, s>>>(T, n, code, w, h0, P, counts);
}
void launch_hist_keys(const uint64_t *K, uint32_t m, const Shifts &sh, uint32_t *counts, uint32_t grid,
                      hipStream_t s) {
    k_hist_keys<<<grid, 256, 0, s>>>(K, m, sh, counts);
}
void launch_sort_pass(const SortArgs &a, hipStream_t s) {
    const uint32_t ntiles = (a.m + SORT_TILE - 1) / SORT_TILE;
    if (a.from_t)
        k_sort_pass<SORT_BLOCK, SORT_ITEMS, true><<<ntiles, SORT_BLOCK, 0, s>>>(
            a.kin, a.vin, a.kout, a.vout, a.m, a.shift, a.dbase, a.status, a.counter, a.epoch, ntiles, a.T, a.code,
            a.w, a.h0);
    else
        k_sort_pass<SORT_BLOCK, SORT_ITEMS, false><<<ntiles, SORT_BLOCK, 0, s>>>(
            a.kin, a.vin, a.kout, a.vout, a.m, a.shift, a.dbase, a.status, a.counter, a.epoch, ntiles, a.T, a.code,
            a.w, a.h0);
}
void launch_regroup(const RegroupArgs &a, hipStream_t s) {
    const uint32_t ntiles = (a.m + RG_TILE - 1) / RG_TILE;
    k_regroup<RG_BLOCK, RG_ITEMS><<<ntiles, RG_BLOCK, 0, s>>>(a.K, a.V, a.pos, a.m, a.kshift, a.SA, a.ISA, a.L, a.T,
                                                             a.pos_out, a.grp_out, a.m_out, a.st_sum, a.st_max,
                                                             a.counter, a.epoch, ntiles);
}
void launch_round_keys(const uint32_t *pos, const uint32_t *grp, uint32_t m, const uint32_t *SA, const uint32_t *ISA,
                       uint32_t n, uint32_t h, uint64_t *K, uint32_t *V, unsigned long long *orand, hipStream_t s) {
    uint32_t grid = (m + 255) / 256;
    if (grid > 2048) grid = 2048;
    k_round_keys<<<grid, 256, 0, s>>>(pos, grp, m, SA, ISA, n, h, K, V, orand);
}
void launch_bwt(const uint8_t *T, const uint8_t *L, const uint32_t *ISA, uint8_t *U, uint32_t n, hipStream_t s) {
    k_bwt<<<(n + 1023) / 1024, 256, 0, s>>>(T, L, ISA, U, n);
}
void launch_samples(const uint32_t *ISA, uint32_t r, uint32_t count, int32_t *I, hipStream_t s) {
    k_samples<<<(count + 255) / 256, 256, 0, s>>>(ISA, r, count, I);
}
void launch_lf_pass(const uint8_t *L, uint32_t n, const uint32_t *cbase, uint32_t *lf, uint64_t *status,
                    uint32_t *counter, uint32_t epoch, hipStream_t s) {
    const uint32_t ntiles = (n + LF_TILE - 1) / LF_TILE;
    k_lf_pass<LF_BLOCK, LF_ITEMS><<<ntiles, LF_BLOCK, 0, s>>>(L, n, cbase, lf, status, counter, epoch, ntiles);
}
void launch_walk_count(const uint32_t *lf, uint32_t n, uint32_t P, uint32_t dshift, uint32_t nreg, uint32_t S,
                       const uint32_t *cbase, uint64_t *link, uint32_t *len, uint8_t *stage, uint32_t cap,
                       uint32_t *cont, hipStream_t s) {
    k_walk_count<<<(S + 1 + 255) / 256, 256, 0, s>>>(lf, n, P, dshift, nreg, S, cbase, link, len, stage, cap, cont);
}
void launch_rank_jump(const uint64_t *in, uint64_t *out, uint32_t S, hipStream_t s) {
    k_rank_jump<<<(S + 1 + 255) / 256, 256, 0, s>>>(in, out, S);
}
void launch_walk_emit(const uint32_t *lf, uint32_t n, uint32_t P, uint32_t S, const uint64_t *link,
                      const uint32_t *len, const uint32_t *cbase, const uint8_t *stage, uint32_t cap,
                      const uint32_t *cont, uint8_t *out, hipStream_t s) {
    k_walk_emit<<<(S + 255) / 256, 256, 0, s>>>(lf, n, P, S, link, len, cbase, stage, cap, cont, out);
}
void launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s) {
    size_t n16 = bytes / 16;
    if (n16) k_copy16<<<2048, 256, 0, s>>>((uint4 *)dst, (const uint4 *)src, n16);
}
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (n) k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
}

}  // namespace kbwt
