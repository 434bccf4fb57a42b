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

__global__ __launch_bounds__(256) void k_byte_hist(const uint8_t *__restrict__ T, uint32_t n,
                                                   uint32_t *__restrict__ partial) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 1024; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += stride) atomicAdd(&h[wave][T[e]], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < 256; b += 256)
        partial[(size_t)blockIdx.x * 256 + b] = h[0][b] + h[1][b] + h[2][b] + h[3][b];
}

// Sums nparts histograms of `bins` words per part (part stride `pstride` words) for each of gridDim.x histogram
// sets (set stride `sstride`); writes totals (if tot) and exclusive scans + base (if excl). One block per set.
__global__ __launch_bounds__(256) void k_scan_digits(const uint32_t *__restrict__ partial, uint32_t nparts,
                                                     uint32_t pstride, uint32_t sstride, uint32_t *__restrict__ tot,
                                                     uint32_t *__restrict__ excl, uint32_t base) {
    __shared__ uint32_t scratch[8];
    const uint32_t set = blockIdx.x, b = threadIdx.x;
    uint32_t s = 0;
    for (uint32_t p = 0; p < nparts; p++) s += partial[(size_t)p * pstride + (size_t)set * sstride + b];
    uint32_t total;
    uint32_t ex = block_excl_sum<256>(s, scratch, &total);
    if (tot) tot[set * 256 + b] = s;
    if (excl) excl[set * 256 + b] = base + ex;
}

__device__ __forceinline__ uint64_t build_key(const uint8_t *c, uint32_t h0, uint32_t w) {
    uint64_t k = 0;
    for (uint32_t b = 0; b < h0; b++) k = (k << w) | c[b];
    return k;
}

// Digit histograms of the first keys (h0 symbols of w bits, zero past the end), P digits of 8 bits.
__global__ __launch_bounds__(256) void k_hist_T(const uint8_t *__restrict__ T, uint32_t n,
                                                const uint8_t *__restrict__ code, uint32_t w, uint32_t h0,
                                                uint32_t P, uint32_t *__restrict__ partial) {
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
            uint64_t key = build_key(s_c + j, h0, w);
            for (uint32_t p = 0; p < P; p++) atomicAdd(&h[p][(uint32_t)(key >> (8 * p)) & 255u], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < MAXP * 256; i += 256) partial[(size_t)blockIdx.x * MAXP * 256 + i] = (&h[0][0])[i];
}

// Digit histograms of stored keys at the given shifts.
__global__ __launch_bounds__(256) void k_hist_keys(const uint64_t *__restrict__ K, uint32_t m, Shifts sh,
                                                   uint32_t *__restrict__ partial) {
    __shared__ uint32_t h[MAXP][256];
    for (int i = threadIdx.x; i < MAXP * 256; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < m; e += stride) {
        uint64_t key = K[e];
        for (uint32_t p = 0; p < sh.P; p++) atomicAdd(&h[p][(uint32_t)(key >> sh.s[p]) & 255u], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < MAXP * 256; i += 256) partial[(size_t)blockIdx.x * MAXP * 256 + i] = (&h[0][0])[i];
}

// ------------------------------------------------------------------ onesweep radix pass

// Stable pass over m (key, value) pairs on the 8-bit digit at `shift`. FROM_T: keys are built from T (value = index).
// dbase[d]: global start of digit d. status: ntiles*256 lookback words; counter: tile-id counter.
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
    __shared__ uint32_t s_wcnt[W][256];
    __shared__ uint32_t s_tstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_scratch[16];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_key[TILE];
    __shared__ uint32_t s_val[TILE];
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
        uint8_t *s_c = reinterpret_cast<uint8_t *>(s_val);   // TILE*4 bytes >= TILE + 64 codes
        for (uint32_t j = tid; j < (uint32_t)TILE + h0; j += BLOCK) {
            size_t e = base + j;
            s_c[j] = e < m ? s_code[T[e]] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint32_t local = wave * WAVE * ITEMS + k * WAVE + lane;
            key[k] = build_key(s_c + local, h0, w);
            val[k] = (uint32_t)(base + local);
        }
        __syncthreads();   // s_val is reused below
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

#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t e = base + wave * WAVE * ITEMS + k * WAVE + lane;
        if (e < m) {
            uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
            uint32_t l = s_tstart[d] + s_wcnt[wave][d] + rk[k];
            s_key[l] = key[k];
            s_val[l] = val[k];
        }
    }
    __syncthreads();
    const uint32_t cnt = (uint32_t)((m - base) < (size_t)TILE ? (m - base) : (size_t)TILE);
    for (uint32_t j = tid; j < cnt; j += BLOCK) {
        uint64_t kk = s_key[j];
        uint32_t d = (uint32_t)(kk >> shift) & 255u;
        uint32_t g = s_gbase[d] + j;
        kout[g] = kk;
        vout[g] = s_val[j];
    }
}

// ------------------------------------------------------------------ regroup (group heads, ISA, compaction)

// Sorted keys K[0..m) with values V (suffixes). Row a goes to SA position pos[a] (pos == nullptr: a). Writes
// SA[pos[a]] = V[a] (when SA != nullptr), ISA[V[a]] = head of its group (SA position of the group's first row) and
// compacts the rows of groups of two or more into (pos_out, grp_out); the last tile writes their count to *m_out.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_regroup(const uint64_t *__restrict__ K, const uint32_t *__restrict__ V,
                                                   const uint32_t *__restrict__ pos, uint32_t m,
                                                   uint32_t *__restrict__ SA, uint32_t *__restrict__ ISA,
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
    uint32_t run_sum = 0, run_max = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        size_t a = base + (size_t)k * BLOCK + tid;
        bool valid = a < m;
        bool flag = false, keep = false;
        uint32_t p = 0;
        if (valid) {
            uint64_t kc = K[a];
            flag = (a == 0) || (K[a - 1] != kc);
            bool flagn = (a + 1 >= m) || (K[a + 1] != kc);
            keep = !(flag && flagn);
            p = pos ? pos[a] : (uint32_t)a;
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
// first). orand[0] |= keys, orand[1] &= keys (which digits vary).
__global__ __launch_bounds__(256) void k_round_keys(const uint32_t *__restrict__ pos, const uint32_t *__restrict__ grp,
                                                    uint32_t m, const uint32_t *__restrict__ SA,
                                                    const uint32_t *__restrict__ ISA, uint32_t n, uint32_t h,
                                                    uint64_t *__restrict__ K, uint32_t *__restrict__ V,
                                                    unsigned long long *__restrict__ orand) {
    size_t a = (size_t)blockIdx.x * 256 + threadIdx.x;
    uint64_t ko = 0, ka = ~0ull;
    if (a < m) {
        uint32_t i = SA[pos[a]];
        uint64_t j = (uint64_t)i + h;
        uint32_t sec = j < n ? ISA[j] + h : n - 1u - i;
        uint64_t key = ((uint64_t)grp[a] << 32) | sec;
        K[a] = key;
        V[a] = i;
        ko = key; ka = key;
    }
    for (int o = 1; o < WAVE; o <<= 1) {
        ko |= __shfl_xor(ko, o, WAVE);
        ka &= __shfl_xor(ka, o, WAVE);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        atomicOr(&orand[0], (unsigned long long)ko);
        atomicAnd(&orand[1], (unsigned long long)ka);
    }
}

// ------------------------------------------------------------------ BWT emission

// U[0] = T[n-1]; row p with SA[p] = i != 0 goes to U[p+1] before the primary row p0 = ISA[0], to U[p] after it.
__global__ __launch_bounds__(256) void k_bwt(const uint8_t *__restrict__ T, const uint32_t *__restrict__ SA,
                                             const uint32_t *__restrict__ ISA, uint8_t *__restrict__ U, uint32_t n) {
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const uint32_t p0 = ISA[0];
    if (p == 0) U[0] = T[n - 1];
    uint32_t i = SA[p];
    size_t o = p < p0 ? p + 1 : p;
    if (i != 0 && o < n) U[o] = T[i - 1];
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

// Splitter s: row s*D for s < nreg, row P for s == nreg (only when P % D != 0). Walk LF from it to the next
// splitter row; link[s] = (steps << 32) | next splitter. The marker row P ends the list: link = (0, S).
__global__ __launch_bounds__(256) void k_walk_count(const uint32_t *__restrict__ lf, uint32_t n, uint32_t P,
                                                    uint32_t dshift, uint32_t nreg, uint32_t S,
                                                    uint64_t *__restrict__ link, uint32_t *__restrict__ len) {
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s > S) return;
    if (s == S) { link[S] = ((uint64_t)0 << 32) | S; return; }
    const uint32_t dmask = (1u << dshift) - 1u;
    const uint32_t sidP = (P & dmask) == 0 ? (P >> dshift) : nreg;
    uint32_t row = s < nreg ? (uint32_t)(s << dshift) : P;
    if (row == P) { link[s] = ((uint64_t)0 << 32) | S; len[s] = 0; return; }
    uint32_t steps = 0;
    do {
        uint32_t q = row < P ? row : row - 1u;
        row = lf[q];
        steps++;
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

// Walk every sublist again, writing its bytes: the splitter's suffix starts at text position acc = link >> 32; the
// byte of row `row` is T[pos-1], pos decreasing by one per LF step. The byte of a row is the c with
// C[c] <= LF(row) < C[c+1].
__global__ __launch_bounds__(256) void k_walk_emit(const uint32_t *__restrict__ lf, uint32_t n, uint32_t P,
                                                   uint32_t dshift, uint32_t nreg, uint32_t S,
                                                   const uint64_t *__restrict__ link, const uint32_t *__restrict__ len,
                                                   const uint32_t *__restrict__ cbase, uint8_t *__restrict__ out) {
    __shared__ uint32_t s_c[257];
    for (int i = threadIdx.x; i < 256; i += 256) s_c[i] = cbase[i];
    if (threadIdx.x == 0) s_c[256] = 0xFFFFFFFFu;
    __syncthreads();
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= S) return;
    uint32_t steps = len[s];
    if (steps == 0) return;
    uint32_t pos = (uint32_t)(link[s] >> 32);
    if (pos > n || steps > pos) return;   // only for a string that is not a BWT; the caller rejects those first
    uint32_t row = s < nreg ? (uint32_t)(s << dshift) : P;
    for (uint32_t t = 0; t < steps; t++) {
        uint32_t q = row < P ? row : row - 1u;
        uint32_t v = lf[q];
        // largest c with C[c] <= v
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t stepc = 128; stepc > 0; stepc >>= 1)
            if (s_c[lo + stepc] <= v) lo += stepc;
        out[pos - 1 - t] = (uint8_t)lo;
        row = v;
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

void launch_byte_hist(const uint8_t *T, uint32_t n, uint32_t *partial, uint32_t grid, hipStream_t s) {
    k_byte_hist<<<grid, 256, 0, s>>>(T, n, partial);
}
void launch_scan_digits(const uint32_t *partial, uint32_t nparts, uint32_t pstride, uint32_t sstride, uint32_t nsets,
                        uint32_t *tot, uint32_t *excl, uint32_t base, hipStream_t s) {
    k_scan_digits<<<nsets, 256, 0, s>>>(partial, nparts, pstride, sstride, tot, excl, base);
}
void launch_hist_T(const uint8_t *T, uint32_t n, const uint8_t *code, uint32_t w, uint32_t h0, uint32_t P,
                   uint32_t *partial, uint32_t grid, hipStream_t s) {
    k_hist_T<<<grid, 256, 0, s>>>(T, n, code, w, h0, P, partial);
}
void launch_hist_keys(const uint64_t *K, uint32_t m, const Shifts &sh, uint32_t *partial, uint32_t grid,
                      hipStream_t s) {
    k_hist_keys<<<grid, 256, 0, s>>>(K, m, sh, partial);
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
void launch_regroup(const uint64_t *K, const uint32_t *V, const uint32_t *pos, uint32_t m, uint32_t *SA,
                    uint32_t *ISA, uint32_t *pos_out, uint32_t *grp_out, uint32_t *m_out, uint64_t *st_sum,
                    uint64_t *st_max, uint32_t *counter, uint32_t epoch, hipStream_t s) {
    const uint32_t ntiles = (m + RG_TILE - 1) / RG_TILE;
    k_regroup<RG_BLOCK, RG_ITEMS><<<ntiles, RG_BLOCK, 0, s>>>(K, V, pos, m, SA, ISA, pos_out, grp_out, m_out, st_sum,
                                                             st_max, counter, epoch, ntiles);
}
void launch_round_keys(const uint32_t *pos, const uint32_t *grp, uint32_t m, const uint32_t *SA, const uint32_t *ISA,
                       uint32_t n, uint32_t h, uint64_t *K, uint32_t *V, unsigned long long *orand, hipStream_t s) {
    k_round_keys<<<(m + 255) / 256, 256, 0, s>>>(pos, grp, m, SA, ISA, n, h, K, V, orand);
}
void launch_bwt(const uint8_t *T, const uint32_t *SA, const uint32_t *ISA, uint8_t *U, uint32_t n, hipStream_t s) {
    k_bwt<<<(n + 255) / 256, 256, 0, s>>>(T, SA, ISA, U, n);
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
                       uint64_t *link, uint32_t *len, hipStream_t s) {
    k_walk_count<<<(S + 1 + 255) / 256, 256, 0, s>>>(lf, n, P, dshift, nreg, S, link, len);
}
void launch_rank_jump(const uint64_t *in, uint64_t *out, uint32_t S, hipStream_t s) {
    k_rank_jump<<<(S + 1 + 255) / 256, 256, 0, s>>>(in, out, S);
}
void launch_walk_emit(const uint32_t *lf, uint32_t n, uint32_t P, uint32_t dshift, uint32_t nreg, uint32_t S,
                      const uint64_t *link, const uint32_t *len, const uint32_t *cbase, uint8_t *out, hipStream_t s) {
    k_walk_emit<<<(S + 255) / 256, 256, 0, s>>>(lf, n, P, dshift, nreg, S, link, len, cbase, out);
}
void launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s) {
    size_t n16 = bytes / 16;
    if (n16) k_copy16<<<2048, 256, 0, s>>>((uint4 *)dst, (const uint4 *)src, n16);
}
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (n) k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
}

}  // namespace kbwt
