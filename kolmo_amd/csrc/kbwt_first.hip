// The end of the first sort (gfx950): buckets sorted in LDS, and the groups they hold.
//
// The first sort orders every suffix by its first key: h0 symbol codes above the BWT byte (section "first keys" in
// kbwt_kernels.hip). The global radix passes sort only the top digits of the key; the rows that share those digits
// form a bucket, a contiguous range of the suffix array. A bucket of at most FG_CAP rows is sorted in LDS by the low
// digits, which ride in the key's low 32 bits together with the BWT byte, and in the same pass its rows are split
// into groups of equal keys: each row writes its suffix (SA), its BWT byte (L) and the pair (suffix, group head) for
// the ISA, and every group of two or more goes to the doubling rounds' lists, as k_regroup does after a full sort.
// Buckets of more rows (rare: long repeats, runs) are listed for the global path, which sorts them by the same low
// digits with the doubling rounds' large-group machinery.
//
// A bucket never crosses a block of a batch, and the rows of a bucket are contiguous, so a workgroup that owns whole
// buckets needs nothing from the others: tile t owns the buckets that start in [t * FG_PITCH, (t + 1) * FG_PITCH).
#include "kbwt_device.h"
#include "kbwt_internal.h"

namespace kbwt {

constexpr int FG_BLOCK = 256, FG_W = FG_BLOCK / WAVE;
constexpr uint32_t FG_ROWS = FG_PITCH + FG_CAP;   // rows a tile holds at most (its buckets end before this)
constexpr int FG_ITEMS = (int)FG_ROWS / FG_BLOCK;
constexpr uint32_t FG_DIRECT = 16;                // buckets of at most this many rows rank each row directly
static_assert(FG_PITCH <= FG_CAP, "a bucket of more than FG_CAP rows must be the last of its tile");
static_assert(FG_ROWS % FG_BLOCK == 0 && FG_ROWS <= 65535, "whole items per thread, 16-bit slots");

// First row of every tile's range: the first bucket start at or after t * FG_PITCH (a row where the top digits, K >>
// hshift, change, or a block starts), so that the tiles' ranges [tstart[t], tstart[t + 1]) hold whole buckets. The
// keys are sorted by their top digits within each block, so a row inside a bucket finds the bucket's end by binary
// search. tstart[ntiles] = N.
__global__ __launch_bounds__(256) void k_first_tiles(const uint64_t *__restrict__ K, uint32_t N, uint32_t hshift,
                                                     const uint16_t *__restrict__ blk,
                                                     const BlockInfo *__restrict__ blocks, uint32_t ntiles,
                                                     uint32_t *__restrict__ tstart) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        tstart[t] = N;
        return;
    }
    const uint32_t a = t * FG_PITCH;   // < N: ntiles = ceil(N / FG_PITCH)
    const BlockInfo bi = blocks[blk ? blk[a] : 0u];
    if (a == bi.off) {
        tstart[t] = a;
        return;
    }
    const uint64_t v = K[a] >> hshift;
    if ((K[a - 1] >> hshift) != v) {
        tstart[t] = a;
        return;
    }
    uint32_t lo = a + 1, hi = bi.off + bi.n;   // first row in (a, end of block) whose top digits exceed v
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if ((K[mid] >> hshift) <= v) lo = mid + 1; else hi = mid;
    }
    tstart[t] = lo;
}

// One tile: the buckets in rows [tstart[t], tstart[t + 1]). K: keys sorted by their top digits (K >> hshift) within
// each block; V: their suffixes. The low 32 bits of a key (lo32 = K & lomask) hold the low digits above the BWT byte,
// and a bucket is sorted by lo32 >> 8. Writes SA[row], L[row] and pairs[row] = (suffix << 32 | head) for every row of
// a bucket of at most FG_CAP rows, appends the groups of two or more to the small (sh, ss) or large (lh, ls) list
// (counters cnt[0], cnt[1]), and appends a bigger bucket, which can only be the tile's last, to (bh, bs) (counters
// cnt[2] buckets, cnt[3] rows) for the global path.
__global__ __launch_bounds__(FG_BLOCK) void k_first_groups(
    const uint64_t *__restrict__ K, const uint32_t *__restrict__ V, const uint32_t *__restrict__ tstart,
    uint32_t hshift, uint32_t lomask, const uint16_t *__restrict__ blk, uint32_t *__restrict__ SA,
    uint64_t *__restrict__ pairs, uint8_t *__restrict__ L, uint32_t *__restrict__ sh, uint32_t *__restrict__ ss,
    uint32_t *__restrict__ lh, uint32_t *__restrict__ ls, uint32_t *__restrict__ bh, uint32_t *__restrict__ bs,
    uint32_t *__restrict__ cnt) {
    __shared__ uint32_t s_key[FG_ROWS], s_val[FG_ROWS];     // rows in tile order: lo32, suffix
    __shared__ uint32_t s_okey[FG_ROWS], s_oval[FG_ROWS];   // each bucket sorted by lo32 >> 8
    __shared__ uint32_t s_flag[FG_ROWS / 32];               // bit j: row j starts a bucket
    __shared__ uint16_t s_wbase[FG_ROWS / 32];              // buckets starting in the words before
    __shared__ uint16_t s_bstart[FG_ROWS];                  // first row of each bucket
    __shared__ uint16_t s_med[FG_ROWS / (FG_DIRECT + 1) + 1];   // buckets the waves sort by radix
    __shared__ uint32_t s_wc[FG_W][256];
    __shared__ uint32_t s_scratch[32];
    __shared__ uint32_t s_nb, s_nmed, s_next, s_base[2];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const uint32_t a0 = tstart[blockIdx.x], a1 = tstart[blockIdx.x + 1];
    if (a0 >= a1) return;
    const uint32_t span = a1 - a0;
    const uint32_t rows = span < FG_ROWS ? span : FG_ROWS;

    // 1. rows into LDS with their bucket-start flags (all loads of a thread in flight together)
    {
        uint64_t kk[FG_ITEMS], kp[FG_ITEMS];
        uint32_t vv[FG_ITEMS];
        bool bb[FG_ITEMS];
#pragma unroll
        for (int k = 0; k < FG_ITEMS; k++) {
            const uint32_t j = (uint32_t)(k * FG_BLOCK + tid);
            const uint32_t a = a0 + (j < rows ? j : 0u);
            kk[k] = K[a];
            kp[k] = K[a > 0 ? a - 1 : 0];
            vv[k] = V[a];
            bb[k] = blk ? blk[a] != blk[a > 0 ? a - 1 : 0] : false;
        }
#pragma unroll
        for (int k = 0; k < FG_ITEMS; k++) {
            const uint32_t j = (uint32_t)(k * FG_BLOCK + tid);
            const bool valid = j < rows;
            const bool f = valid && (j == 0 || (kk[k] >> hshift) != (kp[k] >> hshift) || bb[k]);
            if (valid) {
                s_key[j] = (uint32_t)kk[k] & lomask;
                s_val[j] = vv[k];
            }
            const uint64_t m = __ballot(f);
            if (lane == 0) {
                const uint32_t w0 = (uint32_t)(k * FG_BLOCK + wave * WAVE) / 32u;
                s_flag[w0] = (uint32_t)m;
                s_flag[w0 + 1] = (uint32_t)(m >> 32);
            }
        }
    }
    __syncthreads();

    // 2. bucket starts in order: a prefix count over the flag words
    constexpr int NWORD = (int)FG_ROWS / 32;
    static_assert(NWORD <= FG_BLOCK, "one flag word per thread");
    const uint32_t myw = tid < NWORD ? s_flag[tid] : 0u;
    uint32_t nb;
    const uint32_t wb = block_excl_sum<FG_BLOCK>((uint32_t)__popc(myw), s_scratch, &nb);
    if (tid < NWORD) {
        s_wbase[tid] = (uint16_t)wb;
        uint32_t m = myw, r = wb;
        while (m) {
            const uint32_t b = (uint32_t)__builtin_ctz(m);
            s_bstart[r++] = (uint16_t)(tid * 32 + b);
            m &= m - 1;
        }
    }
    if (tid == 0) {
        s_nb = nb;
        s_nmed = 0;
        s_next = 0;
    }
    __syncthreads();
    // the last bucket ends at the range's end; a bucket of more than FG_CAP rows is the last and goes to the global path
    const uint32_t last = s_bstart[nb - 1];
    const bool big = span - last > FG_CAP;
    const uint32_t send = big ? last : span;   // rows [0, send) belong to buckets sorted here (send <= rows)
    auto bucket_end = [&](uint32_t b) -> uint32_t { return b + 1 < nb ? (uint32_t)s_bstart[b + 1] : span; };

    // 3. buckets of more than FG_DIRECT rows (and at most FG_CAP) into the waves' work list
    for (uint32_t b0 = 0; b0 < nb; b0 += FG_BLOCK) {
        const uint32_t b = b0 + (uint32_t)tid;
        bool med = false;
        if (b < nb) {
            const uint32_t sz = bucket_end(b) - s_bstart[b];
            med = sz > FG_DIRECT && sz <= FG_CAP;
        }
        const uint64_t m = __ballot(med);
        uint32_t wbase_m = 0;
        if (lane == 0 && m) wbase_m = atomicAdd(&s_nmed, (uint32_t)__popcll(m));
        wbase_m = __shfl(wbase_m, 0, WAVE);
        if (med) s_med[wbase_m + count_below(m)] = (uint16_t)b;
    }

    // 4. rows of buckets of at most FG_DIRECT rows: rank = keys below + equal keys before (a row's place in its bucket)
#pragma unroll
    for (int k = 0; k < FG_ITEMS; k++) {
        const uint32_t j = (uint32_t)(k * FG_BLOCK + tid);
        if (j >= send) continue;
        const uint32_t wj = j >> 5;
        const uint32_t b = s_wbase[wj] + (uint32_t)__popc(s_flag[wj] & (0xFFFFFFFFu >> (31 - (j & 31)))) - 1u;
        const uint32_t s = s_bstart[b], e = bucket_end(b);
        if (e - s > FG_DIRECT) continue;
        const uint32_t key = s_key[j], c = key >> 8;
        uint32_t rk = 0;
        for (uint32_t i = s; i < e; i++) {
            const uint32_t ci = s_key[i] >> 8;
            rk += (ci < c || (ci == c && i < j)) ? 1u : 0u;
        }
        s_okey[s + rk] = key;
        s_oval[s + rk] = s_val[j];
    }
    __syncthreads();

    // 5. the other buckets: a wave takes one at a time and sorts it by a stable LSD radix sort over the 8-bit digits
    // of lo32 >> 8 that vary in it, ping-ponging between the input and output arrays (same rows)
    const uint32_t nmed = s_nmed;
    for (;;) {
        uint32_t q = 0;
        if (lane == 0) q = atomicAdd(&s_next, 1u);
        q = __shfl(q, 0, WAVE);
        if (q >= nmed) break;
        const uint32_t b = s_med[q];
        const uint32_t s = s_bstart[b], e = bucket_end(b);
        uint32_t o = 0, an = ~0u;
        for (uint32_t i = s + (uint32_t)lane; i < e; i += WAVE) {
            const uint32_t c = s_key[i] >> 8;
            o |= c;
            an &= c;
        }
        for (int off = 1; off < WAVE; off <<= 1) {
            o |= __shfl_xor(o, off, WAVE);
            an &= __shfl_xor(an, off, WAVE);
        }
        const uint32_t vary = o ^ an;
        uint32_t *kin = s_key, *vin = s_val, *kout = s_okey, *vout = s_oval;
        // The wave's 256 counters are shared by its lanes, which read what other lanes wrote with no barrier in
        // between (the lanes run in step). Every access is an atomic one at wavefront scope, so that it is a real
        // LDS access in program order: with plain accesses the compiler may hand a lane its own earlier store (the
        // zero) for a later load, and a lane that added nothing in the counting loop then misses the other lanes'
        // adds.
        uint32_t *wc = s_wc[wave];
        auto wc_ld = [&](uint32_t x) { return __hip_atomic_load(&wc[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT); };
        auto wc_st = [&](uint32_t x, uint32_t v) {
            __hip_atomic_store(&wc[x], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        };
        for (int d = 0; d < 3; d++) {
            if (((vary >> (8 * d)) & 255u) == 0) continue;
            const uint32_t sh8 = 8 + 8 * (uint32_t)d;
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < 4; u++) wc_st(lane * 4 + u, 0u);
            __builtin_amdgcn_wave_barrier();
            for (uint32_t i = s + (uint32_t)lane; i < e; i += WAVE)
                __hip_atomic_fetch_add(&wc[(kin[i] >> sh8) & 255u], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            __builtin_amdgcn_wave_barrier();
            {   // exclusive scan of the wave's 256 counters, four per lane
                uint32_t c4[4], t4 = 0;
#pragma unroll
                for (int u = 0; u < 4; u++) { c4[u] = wc_ld(lane * 4 + u); t4 += c4[u]; }
                uint32_t ex = wave_incl_sum(t4) - t4;
#pragma unroll
                for (int u = 0; u < 4; u++) { wc_st(lane * 4 + u, s + ex); ex += c4[u]; }
            }
            __builtin_amdgcn_wave_barrier();
            for (uint32_t i0 = s; i0 < e; i0 += WAVE) {   // chunks in row order: stable
                const uint32_t i = i0 + (uint32_t)lane;
                const bool valid = i < e;
                const uint32_t kv = valid ? kin[i] : 0u, vv = valid ? vin[i] : 0u;
                const uint32_t dg = (kv >> sh8) & 255u;
                const uint64_t peers = match_digit8(dg, __ballot(valid));
                const uint32_t c = wc_ld(dg);   // every lane reads before any lane writes
                __builtin_amdgcn_wave_barrier();
                if (valid) {
                    const uint32_t at = c + count_below(peers);
                    kout[at] = kv;
                    vout[at] = vv;
                    wc_st(dg, c + (uint32_t)__popcll(peers));   // every peer writes the same value
                }
                __builtin_amdgcn_wave_barrier();
            }
            uint32_t *t = kin; kin = kout; kout = t;
            t = vin; vin = vout; vout = t;
        }
        __builtin_amdgcn_wave_barrier();
        if (kin != s_okey)
            for (uint32_t i = s + (uint32_t)lane; i < e; i += WAVE) {
                s_okey[i] = s_key[i];
                s_oval[i] = s_val[i];
            }
    }
    __syncthreads();

    // 6. groups: rows j-1 and j share one when they lie in one bucket and agree in lo32 >> 8. Thread t takes rows
    // t*FG_ITEMS .. t*FG_ITEMS + FG_ITEMS - 1; heads by a block max-scan of group starts.
    const uint32_t r0 = (uint32_t)tid * FG_ITEMS;
    uint32_t gflags = 0, gends = 0, tmax = 0;
#pragma unroll
    for (int k = 0; k < FG_ITEMS; k++) {
        const uint32_t j = r0 + k;
        if (j >= send) continue;
        const bool bstart = (s_flag[j >> 5] >> (j & 31)) & 1u;
        const bool gs = bstart || (s_okey[j] >> 8) != (s_okey[j - 1] >> 8);
        const bool nxs = j + 1 >= send || ((s_flag[(j + 1) >> 5] >> ((j + 1) & 31)) & 1u) ||
                         (s_okey[j + 1] >> 8) != (s_okey[j] >> 8);
        if (gs) { gflags |= 1u << k; tmax = j; }
        else if (nxs) gends |= 1u << k;
    }
    uint32_t ex_s, in_m, tot_s, tot_m;
    block_scan_sum_max<FG_BLOCK>(0u, tmax, ex_s, in_m, tot_s, tot_m, s_scratch);
    // the head before this thread's first row: the last group start of the threads before (row 0 always starts one).
    // s_key and s_val are free now (every bucket's sorted rows are in s_okey, s_oval).
    uint32_t *s_imax = s_key, *s_head = s_val;
    s_imax[tid] = in_m;
    __syncthreads();
    uint32_t head = tid > 0 ? s_imax[tid - 1] : 0u;
    uint32_t ns = 0, nl = 0;
    uint32_t eh[FG_ITEMS], es[FG_ITEMS];
#pragma unroll
    for (int k = 0; k < FG_ITEMS; k++) {
        const uint32_t j = r0 + k;
        eh[k] = 0;
        es[k] = 0;
        if (j >= send) continue;
        if (gflags & (1u << k)) head = j;
        s_head[j] = head;
        if (gends & (1u << k)) {
            eh[k] = a0 + head;
            es[k] = j - head + 1;
            if (es[k] > (uint32_t)SEG_GROUP) nl++; else ns++;
        }
    }
    uint32_t tot;
    const uint32_t exl = block_excl_sum<FG_BLOCK>(ns | (nl << 16), s_scratch, &tot);
    if (tid == 0) {
        s_base[0] = (tot & 0xFFFFu) ? atomicAdd(&cnt[0], tot & 0xFFFFu) : 0u;
        s_base[1] = (tot >> 16) ? atomicAdd(&cnt[1], tot >> 16) : 0u;
        if (big) {
            const uint32_t ib = atomicAdd(&cnt[2], 1u);
            bh[ib] = a0 + last;
            bs[ib] = span - last;
            atomicAdd(&cnt[3], span - last);
        }
    }
    __syncthreads();
    uint32_t at_s = s_base[0] + (exl & 0xFFFFu), at_l = s_base[1] + (exl >> 16);
#pragma unroll
    for (int k = 0; k < FG_ITEMS; k++) {
        if (es[k] > (uint32_t)SEG_GROUP) { lh[at_l] = eh[k]; ls[at_l] = es[k]; at_l++; }
        else if (es[k]) { sh[at_s] = eh[k]; ss[at_s] = es[k]; at_s++; }
    }
    // 7. the rows in row order: suffix, BWT byte, (suffix, head) pair
#pragma unroll
    for (int k = 0; k < FG_ITEMS; k++) {
        const uint32_t j = (uint32_t)(k * FG_BLOCK + tid);
        if (j >= send) continue;
        const uint32_t v = s_oval[j];
        SA[a0 + j] = v;
        L[a0 + j] = (uint8_t)s_okey[j];
        pairs[a0 + j] = ((uint64_t)v << 32) | (a0 + s_head[j]);
    }
}

// Rows of the buckets the tiles left to the global path, as one element list for the radix passes and k_regroup:
// buckets g = 0..G-1 in increasing row order, first row bh[g], list offset lo[g] (prefix sums of their sizes), m rows.
// Element a of bucket g is row pos[a] = bh[g] + (a - lo[g]); its key is (g << 32) | lo32 of the row's first key, its
// value the row's suffix. orand[0] |= keys, orand[1] &= keys over bits 8 and up (which digits vary).
__global__ __launch_bounds__(256) void k_big_elems(const uint32_t *__restrict__ bh, const uint32_t *__restrict__ lo,
                                                   uint32_t G, uint32_t m, const uint64_t *__restrict__ K,
                                                   const uint32_t *__restrict__ V, uint32_t lomask,
                                                   uint32_t *__restrict__ pos, uint64_t *__restrict__ K2,
                                                   uint32_t *__restrict__ V2, unsigned long long *__restrict__ orand) {
    __shared__ uint64_t s_o[4], s_a[4];
    uint64_t ko = 0, ka = ~0ull;
    for (size_t a = (size_t)blockIdx.x * 256 + threadIdx.x; a < m; a += (size_t)gridDim.x * 256) {
        uint32_t l = 0, h = G;   // last g with lo[g] <= a
        while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (lo[mid] <= (uint32_t)a) l = mid; else h = mid;
        }
        const uint32_t row = bh[l] + ((uint32_t)a - lo[l]);
        const uint64_t key = ((uint64_t)l << 32) | ((uint32_t)K[row] & lomask);
        pos[a] = row;
        K2[a] = key;
        V2[a] = V[row];
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
        atomicOr(&orand[0], (unsigned long long)(ko & ~0xFFull));
        atomicAnd(&orand[1], (unsigned long long)(ka | 0xFFull));
    }
}

#ifndef KBWT_CPU_EMU   // the launch wrappers need hipcc; a CPU build of the kernels above leaves them out
void launch_first_tiles(const uint64_t *K, uint32_t N, uint32_t hshift, const uint16_t *blk, const BlockInfo *blocks,
                        uint32_t ntiles, uint32_t *tstart, hipStream_t s) {
    k_first_tiles<<<(ntiles + 1 + 255) / 256, 256, 0, s>>>(K, N, hshift, blk, blocks, ntiles, tstart);
}

void launch_first_groups(const FirstGroupsArgs &a, uint32_t ntiles, hipStream_t s) {
    if (!ntiles) return;
    k_first_groups<<<ntiles, FG_BLOCK, 0, s>>>(a.K, a.V, a.tstart, a.hshift, a.lomask, a.blk, a.SA, a.pairs, a.L,
                                               a.sh, a.ss, a.lh, a.ls, a.bh, a.bs, a.cnt);
}

void launch_big_elems(const uint32_t *bh, const uint32_t *lo, uint32_t G, uint32_t m, const uint64_t *K,
                      const uint32_t *V, uint32_t lomask, uint32_t *pos, uint64_t *K2, uint32_t *V2,
                      unsigned long long *orand, hipStream_t s) {
    uint32_t grid = (m + 255) / 256;
    if (grid > 4096) grid = 4096;
    if (grid) k_big_elems<<<grid, 256, 0, s>>>(bh, lo, G, m, K, V, lomask, pos, K2, V2, orand);
}

#endif

}  // namespace kbwt
