// Kernels of the forward suffix sort (prefix doubling), the BWT emission and the inverse BWT, for gfx950.
//
// Every kernel works on a batch of blocks packed back to back (one block is the batch of one). Tiled kernels map a
// tile id to its block through the block's first tile (Tiling); lookback chains never cross a block.
//
// Forward, per batch:
//   1. k_byte_hist + k_scan_digits: byte histogram of every block; the host turns each into a dense order-preserving
//      symbol code of w = ceil(log2 sigma) bits and picks h0 symbols per first key (h0*w <= 56 bits).
//   2. k_hist_gram (and k_hist_T for 5-bit codes) + k_scan_digits: the 8-bit digit histograms of every first key of
//      every block, in one read of T (none for 8-bit codes: their byte histogram serves), from one count per position
//      of the codes' m-grams; exclusive scans give each block's digit buckets (onesweep).
//   3. k_sort_pass: one stable LSD pass per digit, segmented by block. The first pass builds the keys from T: the
//      symbol codes above the BWT byte T[i-1]. A tile of BLOCK*ITEMS elements is ranked by wave64 ballot multisplit,
//      its digit offsets come from a decoupled lookback over earlier tiles of its block, and it leaves through LDS
//      so that equal digits go out as contiguous runs.
//   4. k_regroup_carry + k_regroup: group heads (rank = SA position of the group's first row) by a max-scan inside
//      each tile, whose carry-in (the head of the group holding the tile's first row) comes from a binary search over
//      the sorted keys, so tiles need nothing from each other; ISA[suffix] = head (in the first round through pairs
//      partitioned by ISA window, whose stretches of ISA are assembled in LDS or, for wide windows, scattered run by
//      run), the row's BWT byte, and the groups of two or more appended to the next round's group lists.
//   5. per doubling round with offset h: groups of at most SEG_GROUP rows sort inside one workgroup each
//      (kbwt_rounds.hip); for the rows of larger groups k_round_keys gathers (the group's index in the head-sorted
//      list, ISA[i+h] + h, or the distance to the block's end when i+h lies past it), the same radix passes sort them
//      (one flat range: the group indexes already keep blocks apart), and k_regroup splits those groups.
//   6. k_bwt + k_samples: U in libsais' convention per block and I[k] = 1 + rank of suffix k*r.
// Inverse, per batch of BWTs L with primary indexes P (the end marker's row in the n+1-row matrix of each block):
//   1. byte histograms -> C[c] = first row of byte c, as a global row.
//   2. k_lf_pass: LF(row) = C[c] + occ(c, row) for every row, a stable multisplit ranking with lookback.
//   3. k_lf_pairs: per row, the row two LF steps on and the first step's byte (the second step's byte follows from
//      the landing row and C), so a walk takes two steps per load.
//   4. k_walk_count: from every splitter row (row % D == 0, and P) follow LF to the next splitter, recording the
//      sublist length and staging its first bytes; k_rank_jump: synchronous pointer jumping gives each splitter's
//      text position; k_walk_emit: each sublist's bytes land at their positions.
#include <type_traits>

#include "kbwt_device.h"
#include "kbwt_internal.h"

namespace kbwt {

// ------------------------------------------------------------------ tiles and blocks

// The block b with first[b] <= t < first[b + 1].
__device__ __forceinline__ uint32_t tile_block(const uint32_t *first, uint32_t B, uint32_t t) {
    uint32_t lo = 0, hi = B;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Smallest s with 256 << s >= n: the width of a block's ISA windows (at most 256 windows per block).
__device__ __forceinline__ uint32_t window_shift(uint32_t n) { return n > 256 ? 24u - (uint32_t)__clz(n - 1) : 0u; }

struct TileRange {
    uint32_t b;        // block
    uint32_t tb;       // tile within the block
    size_t beg, end;   // element range of the tile
    size_t bend;       // end of the block
};

__device__ __forceinline__ TileRange tile_range(const Tiling &tl, uint32_t t, uint32_t tile, size_t m) {
    TileRange r;
    if (tl.first) {
        r.b = tile_block(tl.first, tl.B, t);
        r.tb = t - tl.first[r.b];
        const BlockInfo bi = tl.blocks[r.b];
        r.beg = (size_t)bi.off + (size_t)r.tb * tile;
        r.bend = (size_t)bi.off + bi.n;
    } else {
        r.b = 0;
        r.tb = t;
        r.beg = (size_t)t * tile;
        r.bend = m;
    }
    r.end = r.beg + tile < r.bend ? r.beg + tile : r.bend;
    return r;
}

// put(j, T[beg + j]) for every j < cnt, with byte 0 where beg + j >= lim (past the block). The aligned middle of
// the range moves in 16-byte loads, one or a few per thread and all in flight together; the unaligned head and
// tail (under 16 bytes each) and the part past lim go byte by byte.
template <int BLOCK, class F>
__device__ __forceinline__ void stage_bytes(const uint8_t *__restrict__ T, size_t beg, uint32_t cnt, size_t lim,
                                            F put) {
    const int tid = threadIdx.x;
    const size_t vend = beg + cnt < lim ? beg + cnt : (lim > beg ? lim : beg);   // end of the real bytes
    const size_t mis = (size_t)(reinterpret_cast<uintptr_t>(T + beg) & 15u);
    const size_t ab = beg + (mis ? 16 - mis : 0) < vend ? beg + (mis ? 16 - mis : 0) : vend;
    const size_t ae = ab + ((vend - ab) & ~(size_t)15);
    const uint4 *vp = reinterpret_cast<const uint4 *>(T + ab);
    const uint32_t nv = (uint32_t)((ae - ab) >> 4), o = (uint32_t)(ab - beg);
    for (uint32_t q = tid; q < nv; q += BLOCK) {
        const uint4 v = vp[q];
        const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; k++) put(o + 16 * q + k, (uint8_t)(wd[k >> 2] >> (8 * (k & 3))));
    }
    for (uint32_t j = tid; j < o; j += BLOCK) put(j, T[beg + j]);
    const uint32_t te = (uint32_t)(vend - beg);
    for (uint32_t j = (uint32_t)(ae - beg) + tid; j < te; j += BLOCK) put(j, T[beg + j]);
    for (uint32_t j = te + tid; j < cnt; j += BLOCK) put(j, (uint8_t)0);
}

// ------------------------------------------------------------------ histograms
//
// Histogram kernels accumulate in LDS and add their non-zero bins to a global `counts` array, one atomic per bin and
// block; k_scan_digits turns counts into exclusive scans and clears them again, so counts is all zero between uses.
// Workgroup g walks a contiguous run of tiles and flushes whenever the block changes.

__global__ __launch_bounds__(256) void k_byte_hist(const uint8_t *__restrict__ T, Tiling tl, uint32_t ntiles,
                                                   uint32_t *__restrict__ counts) {
    // BH_COPIES sub-histograms, lane l counting into copy l % BH_COPIES: equal bytes in one wave collide only within a
    // copy (text is skewed; one histogram would serialise its frequent bytes)
    constexpr int BH_COPIES = 16;
    __shared__ uint32_t h[BH_COPIES][256];
    const int tid = threadIdx.x, cp = tid % BH_COPIES;
    for (int i = tid; i < BH_COPIES * 256; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    int cur = -1;
    auto flush = [&] {
        __syncthreads();
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < BH_COPIES; k++) { c += h[k][tid]; h[k][tid] = 0; }
        if (c) atomicAdd(&counts[(size_t)cur * 256 + tid], c);
        __syncthreads();
    };
    // A tile is at most one 16-byte vector per thread plus single head and tail bytes: each thread loads its share of
    // the next tile before it counts the current one, so the loads' latency overlaps the atomics.
    static_assert(HIST_TILE == 256 * 16, "one vector per thread and tile");
    struct Share {
        uint4 v;
        uint32_t nv, head, tail, b;   // bytes of v that count; head/tail byte + 256 when present; block
    };
    auto fetch = [&](uint32_t t) {
        Share sh{make_uint4(0, 0, 0, 0), 0u, 0u, 0u, 0u};
        const TileRange r = tile_range(tl, t, HIST_TILE, 0);
        sh.b = r.b;
        const size_t mis = (size_t)(reinterpret_cast<uintptr_t>(T + r.beg) & 15u);
        const size_t ab = r.beg + (mis ? 16 - mis : 0) < r.end ? r.beg + (mis ? 16 - mis : 0) : r.end;
        const size_t ae = ab + ((r.end - ab) & ~(size_t)15);
        if ((size_t)tid < ((ae - ab) >> 4)) {
            sh.v = reinterpret_cast<const uint4 *>(T + ab)[tid];   // 16-byte aligned address
            sh.nv = 16;
        }
        if ((size_t)tid < ab - r.beg) sh.head = 256u + T[r.beg + tid];
        if ((size_t)tid < r.end - ae) sh.tail = 256u + T[ae + tid];
        return sh;
    };
    Share sh = t0 < t1 ? fetch(t0) : Share{};
    for (uint32_t t = t0; t < t1; t++) {
        const Share cur_sh = sh;
        if (t + 1 < t1) sh = fetch(t + 1);
        if ((int)cur_sh.b != cur) {
            if (cur >= 0) flush();
            cur = (int)cur_sh.b;
        }
        if (cur_sh.nv) {
            const uint32_t wd[4] = {cur_sh.v.x, cur_sh.v.y, cur_sh.v.z, cur_sh.v.w};
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int k = 0; k < 4; k++) atomicAdd(&h[cp][(wd[q] >> (8 * k)) & 255u], 1u);
        }
        if (cur_sh.head) atomicAdd(&h[cp][cur_sh.head - 256u], 1u);
        if (cur_sh.tail) atomicAdd(&h[cp][cur_sh.tail - 256u], 1u);
    }
    if (cur >= 0) flush();
}


// One workgroup per histogram set (set = block * sets_per_block + p): tot = counts, excl = base + exclusive scan of
// counts, and counts is cleared. base: 0 (mode 0), the block's first packed index (1), its first row + 1 (2).
__global__ __launch_bounds__(256) void k_scan_digits(uint32_t *__restrict__ counts, uint32_t sets_per_block,
                                                     const BlockInfo *__restrict__ blocks, int base_mode,
                                                     uint32_t *__restrict__ tot, uint32_t *__restrict__ excl) {
    __shared__ uint32_t scratch[8];
    const uint32_t set = blockIdx.x, d = threadIdx.x;
    const uint32_t b = set / sets_per_block;
    const uint32_t base = base_mode == 0 ? 0u : (base_mode == 1 ? blocks[b].off : blocks[b].roff + 1u);
    const size_t at = (size_t)set * 256 + d;
    const uint32_t s = counts[at];
    uint32_t total;
    uint32_t ex = block_excl_sum<256>(s, scratch, &total);
    if (tot) tot[at] = s;
    if (excl) excl[at] = base + ex;
    counts[at] = 0;
}

// First key of suffix i: h0 symbol codes of w bits (zero past the block's end) above the 8-bit BWT byte T[i-1] (0 for
// the block's first suffix, the primary row, whose byte is never emitted). c points at the code of T[i].
__device__ __forceinline__ uint64_t build_key(const uint8_t *c, uint8_t bwt_byte, uint32_t h0, uint32_t w) {
    uint64_t k = 0;
    for (uint32_t b = 0; b < h0; b++) k = (k << w) | c[b];
    return (k << 8) | bwt_byte;
}

// Digit histograms of the first keys of every block, P digits of 8 bits above the BWT byte: counts[(b*P + p)*256 + d].
__global__ __launch_bounds__(256) void k_hist_T(const uint8_t *__restrict__ T, Tiling tl, uint32_t ntiles,
                                                const uint8_t *__restrict__ codes, uint32_t P,
                                                uint32_t *__restrict__ counts) {
    // HT_COPIES sub-histogram sets, lane l counting into set l % HT_COPIES (skewed digits collide only within a
    // set), each bin a 16-bit half of a word: two bins per atomic word keep all sets in 32 KB of LDS. A set sees at
    // most HT_FLUSH tiles / HT_COPIES rows between flushes, so no half overflows.
    constexpr int HT_COPIES = 8, HT_FLUSH = 64;
    static_assert((size_t)HT_FLUSH * HIST_TILE / HT_COPIES < 65536, "16-bit bins");
    // Thread t keys the RUN consecutive rows from t * RUN, sliding its key one code per row. The codes sit in LDS
    // with 4 spare bytes after every 16, so the runs of a wave's lanes start in different banks.
    constexpr int RUN = HIST_TILE / 256;
    static_assert(RUN == 16, "one 16-byte chunk per thread");
    __shared__ uint32_t h[HT_COPIES][MAXP][128];
    __shared__ uint8_t s_code[256];
    __shared__ uint8_t s_c[(HIST_TILE + 64) / 16 * 20];
    auto at = [](uint32_t j) { return j + ((j >> 4) << 2); };
    const int tid = threadIdx.x, cp = tid % HT_COPIES;
    for (int i = tid; i < HT_COPIES * MAXP * 128; i += 256) (&h[0][0][0])[i] = 0;
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    int cur = -1;
    uint32_t w = 0, h0 = 0, since = 0;
    auto flush = [&](uint32_t b) {   // every thread, after a barrier
        for (uint32_t i = tid; i < P * 256; i += 256) {
            const uint32_t p = i >> 8, d = i & 255u;
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < HT_COPIES; k++) c += (h[k][p][d >> 1] >> (16 * (d & 1))) & 0xFFFFu;
            if (c) atomicAdd(&counts[(size_t)b * P * 256 + i], c);
        }
        __syncthreads();
        for (int i = tid; i < HT_COPIES * MAXP * 128; i += 256) (&h[0][0][0])[i] = 0;
    };
    for (uint32_t t = t0; t < t1; t++) {
        TileRange r = tile_range(tl, t, HIST_TILE, 0);
        if (gram_block(tl.blocks[r.b].w)) continue;   // k_hist_gram's block
        __syncthreads();
        if ((int)r.b != cur || since == (uint32_t)HT_FLUSH) {
            if (cur >= 0) flush((uint32_t)cur);
            since = 0;
            if ((int)r.b != cur) {
                cur = (int)r.b;
                s_code[tid] = codes[(size_t)r.b * 256 + tid];
                w = tl.blocks[r.b].w;
                h0 = tl.blocks[r.b].h0;
            }
            __syncthreads();
        }
        since++;
        stage_bytes<256>(T, r.beg, HIST_TILE + h0, r.bend,
                         [&](uint32_t j, uint8_t t) { s_c[at(j)] = r.beg + j < r.bend ? s_code[t] : 0; });
        __syncthreads();
        const uint32_t j0 = (uint32_t)tid * RUN;
        const uint32_t nrow = r.end - r.beg;
        if (j0 < nrow) {
            const uint64_t kmask = (1ull << (h0 * w)) - 1ull;   // h0 * w <= 56
            uint64_t k = 0;
            for (uint32_t c = 0; c + 1 < h0; c++) k = (k << w) | s_c[at(j0 + c)];
            const uint32_t nr = nrow - j0 < (uint32_t)RUN ? nrow - j0 : (uint32_t)RUN;
            for (uint32_t i = 0; i < nr; i++) {
                k = ((k << w) | s_c[at(j0 + i + h0 - 1)]) & kmask;   // the codes of rows j0 + i .. j0 + i + h0 - 1
                for (uint32_t p = 0; p < P; p++) {
                    const uint32_t d = (uint32_t)(k >> (8 * p)) & 255u;
                    atomicAdd(&h[cp][p][d >> 1], 1u << (16 * (d & 1)));
                }
            }
        }
    }
    __syncthreads();
    if (cur >= 0) flush((uint32_t)cur);
}

// m-gram digit histograms. Digit p of the first key of suffix i is bits [8p, 8p + 8) of its h0 codes (the code of
// T[i + h0 - 1] lowest, so code c is the symbol at i + h0 - 1 - c). It overlaps at most m = gram_len(w) consecutive
// codes, from lo = floor(8p / w) up, so it is a bit field of the m-gram G(j) = x_j .. x_{j+m-1} (x the codes, zero
// outside the block, the first one highest) at j = i + o, o = h0 - lo - m, with the gram's codes before position i
// masked off. A block's digit histograms are therefore the histogram M of G(j) over j in [0, n), each gram adding its
// count to the bin it gives in every digit, corrected at the block's ends: the grams G(i + o) that start outside
// [0, n) are added and the grams in [0, n) that belong to no suffix removed (at most |o| <= 56 of each per digit).
// That is one LDS count per position instead of one per digit.
__device__ __forceinline__ uint32_t gram_digit(uint32_t g, uint32_t p, uint32_t w, uint32_t h0, uint32_t m) {
    if (8 * p >= h0 * w) return 0;   // above the key: always zero
    const uint32_t lo = 8 * p / w, s = 8 * p - w * lo, kb = w * (h0 - lo);   // kb: gram bits inside the key
    return ((kb < m * w ? g & ((1u << kb) - 1u) : g) >> s) & 255u;
}

constexpr int GRAM_BLOCK = 1024, GRAM_CHUNK = GRAM_BLOCK * 16;
static_assert(GRAM_UNIT % GRAM_CHUNK == 0, "whole chunks per unit");

// Bytes [p, p + 16) of a block ending at lim and how many of them lie inside it; one aligned vector load when all do
// (the caller keeps p 16-byte aligned then).
struct Raw16 {
    uint4 v;
    uint32_t nv;
};
__device__ __forceinline__ Raw16 load_raw16(const uint8_t *__restrict__ T, size_t p, size_t lim) {
    Raw16 r;
    r.nv = p >= lim ? 0u : (lim - p >= 16 ? 16u : (uint32_t)(lim - p));
    if (r.nv == 16) {
        r.v = *reinterpret_cast<const uint4 *>(T + p);
    } else {
        uint32_t wd[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 16; k++)
            if (k < r.nv) wd[k >> 2] |= (uint32_t)T[p + k] << (8 * (k & 3));
        r.v = make_uint4(wd[0], wd[1], wd[2], wd[3]);
    }
    return r;
}
// The codes of a Raw16 (0 past the block), packed the same way.
__device__ __forceinline__ uint4 codes16(const Raw16 &r, const uint8_t *s_code) {
    const uint32_t wd[4] = {r.v.x, r.v.y, r.v.z, r.v.w};
    uint32_t o[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        uint32_t x = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (4 * q + k < r.nv) x |= (uint32_t)s_code[(wd[q] >> (8 * k)) & 255u] << (8 * k);
        o[q] = x;
    }
    return make_uint4(o[0], o[1], o[2], o[3]);
}
// Count the M-grams starting at chunk positions j0 .. j0 + 15 below cnt: the 16 + 8 codes from j0 in two LDS loads,
// then 16 independent LDS atomics.
template <uint32_t M>
__device__ __forceinline__ void count_grams(const uint8_t *s_c, uint32_t j0, uint32_t cnt, uint32_t w, uint32_t gmask,
                                            uint32_t *h) {
    const uint4 a = *reinterpret_cast<const uint4 *>(s_c + j0);
    const uint2 b = *reinterpret_cast<const uint2 *>(s_c + j0 + 16);
    const uint32_t wd[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
    auto code = [&](uint32_t k) { return (wd[k >> 2] >> (8 * (k & 3))) & 255u; };
    uint32_t g = 0;
#pragma unroll
    for (uint32_t k = 0; k + 1 < M; k++) g = (g << w) | code(k);
    if (j0 + 16 <= cnt) {
#pragma unroll
        for (uint32_t i = 0; i < 16; i++) {
            g = ((g << w) | code(i + M - 1)) & gmask;
            atomicAdd(&h[g], 1u);
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < 16; i++)
            if (j0 + i < cnt) {
                g = ((g << w) | code(i + M - 1)) & gmask;
                atomicAdd(&h[g], 1u);
            }
    }
}

// One workgroup per GRAM_UNIT of a block with gram_block(w): M of the unit in LDS, mapped to the P digit histograms in
// LDS, added to counts[(b*P + p)*256 + d]. The workgroup of the block's first unit adds the end corrections. 8-bit
// codes take one unit per block and no pass over T: their 1-grams are the byte histogram tot. The unit goes through
// LDS in chunks of 16 codes per thread; when the unit starts 16-byte aligned, each thread translates its 16 bytes in
// registers, stores them with one vector write and loads the next chunk's bytes while the current one is counted.
__global__ __launch_bounds__(GRAM_BLOCK) void k_hist_gram(const uint8_t *__restrict__ T, Tiling units,
                                                          const uint8_t *__restrict__ codes,
                                                          const uint32_t *__restrict__ tot, uint32_t P,
                                                          uint32_t *__restrict__ counts) {
    __shared__ uint32_t h[1u << GRAM_MAXBITS];
    __shared__ uint32_t dh[MAXP][256];
    __shared__ uint8_t s_code[256];
    __shared__ __align__(16) uint8_t s_c[GRAM_CHUNK + 16];   // a chunk's codes and the next few
    __shared__ uint8_t s_end[2][64];                         // codes of the block's first and last 64 positions
    const uint32_t tid = threadIdx.x;
    const TileRange r = tile_range(units, blockIdx.x, GRAM_UNIT, 0);
    const BlockInfo bi = units.blocks[r.b];
    const uint32_t w = bi.w, h0 = bi.h0, n = bi.n, m = gram_len(w), nb = 1u << (m * w), gmask = nb - 1u;
    for (uint32_t i = tid; i < nb; i += GRAM_BLOCK) h[i] = 0;
    for (uint32_t i = tid; i < MAXP * 256; i += GRAM_BLOCK) (&dh[0][0])[i] = 0;
    if (tid < 256) s_code[tid] = codes[(size_t)r.b * 256 + tid];
    __syncthreads();
    if (r.tb == 0 && tid < 128) {
        const uint32_t e = tid >> 6, k = tid & 63u;
        const int64_t q = e ? (int64_t)n - 64 + k : (int64_t)k;
        s_end[e][k] = q >= 0 && q < (int64_t)n ? s_code[T[(size_t)bi.off + (size_t)q]] : 0;
    }
    if (w == 8) {   // distinct bytes have distinct codes
        if (tid < 256) {
            const uint32_t c = tot[(size_t)r.b * 256 + tid];
            if (c) h[s_code[tid]] = c;
        }
    } else {
        const uint32_t nch = (uint32_t)((r.end - r.beg + GRAM_CHUNK - 1) / GRAM_CHUNK);
        const bool vec = ((reinterpret_cast<uintptr_t>(T) + r.beg) & 15u) == 0;   // then every chunk starts aligned
        Raw16 v{};                  // this thread's 16 bytes of the chunk to come
        uint32_t la = 0, la_in = 0;   // threads 0-7: one byte past that chunk, and whether it lies in the block
        auto fetch = [&](size_t c0) {
            v = load_raw16(T, c0 + 16 * tid, r.bend);
            if (tid < 8) {
                const size_t q = c0 + GRAM_CHUNK + tid;
                la_in = q < r.bend;
                la = la_in ? T[q] : 0u;
            }
        };
        if (vec) fetch(r.beg);
        for (uint32_t k = 0; k < nch; k++) {
            const size_t c0 = r.beg + (size_t)k * GRAM_CHUNK;
            const uint32_t cnt = (uint32_t)(r.end - c0 < (size_t)GRAM_CHUNK ? r.end - c0 : (size_t)GRAM_CHUNK);
            if (vec) {
                *reinterpret_cast<uint4 *>(s_c + 16 * tid) = codes16(v, s_code);
                if (tid < 8) s_c[GRAM_CHUNK + tid] = la_in ? s_code[la] : 0;
                __syncthreads();
                if (k + 1 < nch) fetch(c0 + GRAM_CHUNK);
            } else {
                stage_bytes<GRAM_BLOCK>(T, c0, cnt + m - 1, r.bend,
                                        [&](uint32_t j, uint8_t t) { s_c[j] = c0 + j < r.bend ? s_code[t] : 0; });
                __syncthreads();
            }
            const uint32_t j0 = 16 * tid;
            if (j0 < cnt) {
                if (m == 2) count_grams<2>(s_c, j0, cnt, w, gmask, h);
                else if (m == 4) count_grams<4>(s_c, j0, cnt, w, gmask, h);
                else count_grams<8>(s_c, j0, cnt, w, gmask, h);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    // The unit's digit histograms. Digit p is bits [s, s + wd) of a gram (s = 8p - w*lo <= 6, wd = min(8, kb - s)), so
    // hist_p[d] is the sum of the entries y = d mod 2^wd of L_s, the table with its low s bits summed out. L_0 = h and
    // L_{s+1}[y] = L_s[2y] + L_s[2y+1] are built in place; every bin is one sum over consecutive words (no atomics).
    // A digit above the key is 0 for every position.
    uint32_t smax = 0;
    for (uint32_t p = 0; p < P; p++)
        if (8 * p < h0 * w) smax = max(smax, 8 * p - w * (8 * p / w));
    if (tid < P && 8 * tid >= h0 * w) dh[tid][0] = w == 8 ? n : (uint32_t)(r.end - r.beg);
    uint32_t len = nb;
    for (uint32_t s = 0;; s++) {
        for (uint32_t i = tid; i < P * 256; i += GRAM_BLOCK) {
            const uint32_t p = i >> 8, d = i & 255u, lo = 8 * p / w;
            if (8 * p >= h0 * w || 8 * p - w * lo != s) continue;
            const uint32_t kb = w * (h0 - lo), wd = kb - s < 8 ? kb - s : 8u;
            if (d >> wd) continue;
            uint32_t sum = 0;
            for (uint32_t y = d; y < len; y += 1u << wd) sum += h[y];
            dh[p][d] = sum;
        }
        if (s == smax) break;
        uint32_t pr[(1u << GRAM_MAXBITS) / 2 / GRAM_BLOCK];
#pragma unroll
        for (uint32_t k = 0; k < (1u << GRAM_MAXBITS) / 2 / GRAM_BLOCK; k++) {
            const uint32_t y = tid + k * GRAM_BLOCK;
            pr[k] = y < len / 2 ? h[2 * y] + h[2 * y + 1] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < (1u << GRAM_MAXBITS) / 2 / GRAM_BLOCK; k++) {
            const uint32_t y = tid + k * GRAM_BLOCK;
            if (y < len / 2) h[y] = pr[k];
        }
        __syncthreads();
        len >>= 1;
    }
    __syncthreads();
    if (r.tb == 0 && tid < P && 8 * tid < h0 * w) {   // end corrections of digit p (counts wrap; the sums are exact)
        const uint32_t p = tid;
        const int64_t o = (int64_t)h0 - (int64_t)(8 * p / w) - (int64_t)m, nn = n;
        // positions read: below h0 <= 56 at the start, from n - 7 at the end
        auto x = [&](int64_t q) -> uint32_t {
            if (q < 0 || q >= nn) return 0;
            return q < 64 ? s_end[0][q] : s_end[1][q - (nn - 64)];
        };
        auto G = [&](int64_t j) {
            uint32_t g = 0;
            for (uint32_t k = 0; k < m; k++) g = (g << w) | x(j + (int64_t)k);
            return g;
        };
        if (o > 0) {   // suffixes i >= n - o read all-zero grams past the end; grams j < o belong to no suffix
            const int64_t e = o < nn ? o : nn;
            for (int64_t j = 0; j < e; j++) atomicSub(&dh[p][gram_digit(G(j), p, w, h0, m)], 1u);
            atomicAdd(&dh[p][0], (uint32_t)e);
        } else if (o < 0) {   // suffixes i < -o read grams starting before 0; grams j >= n + o belong to no suffix
            for (int64_t j = nn + o > 0 ? nn + o : 0; j < nn; j++)
                atomicSub(&dh[p][gram_digit(G(j), p, w, h0, m)], 1u);
            const int64_t e = -o < nn ? -o : nn;
            for (int64_t i = 0; i < e; i++) atomicAdd(&dh[p][gram_digit(G(i + o), p, w, h0, m)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < P * 256; i += GRAM_BLOCK) {
        const uint32_t v = (&dh[0][0])[i];
        if (v) atomicAdd(&counts[(size_t)r.b * P * 256 + i], v);
    }
}

// Digit histograms of stored keys at the given shifts (one flat range).
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

// Stable pass over (key, value) pairs on the 8-bit digit at `shift`, segmented by block when tl.first is set.
// FROM_T: keys are built from T with the block's code table (value = the suffix index); this pass and the keys-only
// one below need not be stable and are not. dbase[b*dstride + d]: start of
// digit d in block b. status: one lookback word per (tile, digit); counter: tile-id counter.
// epoch_arg: the lookback epoch, or with depoch its offset from the device epoch *depoch (a replayed graph reads its
// epoch there).
// Keys-only passes (VALS == false) move (suffix << 32 | head) pairs alone, partitioned by the suffix's window inside
// its block b: digit (suffix - off_b) >> window_shift(n_b). The suffixes of a block are a permutation of its own
// index range, so window d starts at off_b + min(d << window_shift(n_b), n_b) (dbase unused).
template <int BLOCK, int ITEMS, int LB, bool FROM_T, bool VALS = true>
__global__ __launch_bounds__(BLOCK) void k_sort_pass(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                     uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     uint32_t m, uint32_t shift, const uint32_t *__restrict__ dbase,
                                                     uint32_t dstride, Tiling tl, uint64_t *__restrict__ status,
                                                     uint32_t *__restrict__ counter, uint32_t epoch_arg,
                                                     const uint32_t *__restrict__ depoch, uint32_t ntiles,
                                                     const uint8_t *__restrict__ T,
                                                     const uint8_t *__restrict__ codes) {
    const uint32_t epoch = depoch ? *depoch + epoch_arg : epoch_arg;
    constexpr int W = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(BLOCK >= 256, "one thread per digit");
    static_assert(TILE * 8 >= (TILE + 64) / 16 * 40, "LDS staging of raw bytes and codes");
    __shared__ uint32_t s_wcnt[W][256];
    __shared__ uint32_t s_tstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_scratch[16];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_buf[TILE];
    __shared__ uint8_t s_code[FROM_T ? 256 : 1];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    // element of item k of this lane: k-major, lane-minor in a wave's run (stable), or ITEMS consecutive elements
    // per lane in the first pass
    auto item = [&](int k) -> uint32_t {
        return FROM_T ? (uint32_t)(wave * WAVE + lane) * ITEMS + k : (uint32_t)(wave * WAVE * ITEMS + k * WAVE + lane);
    };
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    for (int i = tid; i < W * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const TileRange tr = tile_range(tl, tile, TILE, m);
    const size_t base = tr.beg;
    const uint32_t cnt = tr.end > tr.beg ? (uint32_t)(tr.end - tr.beg) : 0u;   // a tile past the keys holds none
    const uint32_t *db = dbase ? dbase + (size_t)tr.b * dstride : nullptr;
    uint32_t woff = 0, wn = 0, wsh = 0;   // keys-only passes: the block's range and window width
    if (!VALS) {
        woff = tl.blocks[tr.b].off;
        wn = tl.blocks[tr.b].n;
        wsh = window_shift(wn);
    }
    auto digit = [&](uint64_t k) -> uint32_t {
        return VALS ? (uint32_t)(k >> shift) & 255u : (((uint32_t)(k >> 32) - woff) >> wsh) & 255u;
    };

    uint64_t key[ITEMS];
    uint32_t val[ITEMS];
    if (FROM_T) {
        // raw[j] = T[base + j - 1] (the BWT byte of element j), cod[j] = code of T[base + j], zero past the block.
        // The first pass need not be stable (its input order is text order, which nothing depends on), so a lane keys
        // ITEMS consecutive elements, sliding its key one code per element; 4 spare bytes after every 16 put the
        // lanes' runs in different LDS banks.
        const uint32_t w = tl.blocks[tr.b].w, h0 = tl.blocks[tr.b].h0;
        const size_t boff = tl.blocks[tr.b].off;
        if (tid < 256) s_code[tid] = codes[(size_t)tr.b * 256 + tid];
        auto at = [](uint32_t j) { return j + ((j >> 4) << 2); };
        uint8_t *s_raw = reinterpret_cast<uint8_t *>(s_buf);
        uint8_t *s_cod = s_raw + (TILE + 64) / 16 * 20;
        __syncthreads();
        stage_bytes<BLOCK>(T, base, (uint32_t)TILE + h0, tr.bend, [&](uint32_t j, uint8_t t) {
            s_cod[at(j)] = base + j < tr.bend ? s_code[t] : 0;
            if (j + 1 <= (uint32_t)TILE) s_raw[at(j + 1)] = t;
        });
        if (tid == 0) s_raw[0] = base > boff ? T[base - 1] : 0;
        __syncthreads();
        const uint32_t j0 = (wave * WAVE + lane) * ITEMS;
        const uint64_t kmask = (1ull << (h0 * w)) - 1ull;   // h0 * w <= 56
        uint64_t kp = 0;
        for (uint32_t c = 0; c + 1 < h0; c++) kp = (kp << w) | s_cod[at(j0 + c)];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            kp = ((kp << w) | s_cod[at(j0 + k + h0 - 1)]) & kmask;
            key[k] = (kp << 8) | s_raw[at(j0 + k)];
            val[k] = (uint32_t)(base + j0 + k);
        }
        __syncthreads();   // s_buf is reused below
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint32_t local = wave * WAVE * ITEMS + k * WAVE + lane;
            if (local < cnt) { key[k] = kin[base + local]; val[k] = VALS ? vin[base + local] : 0u; }
            else { key[k] = 0; val[k] = 0; }
        }
    }

    // ranking inside each wave in (k, lane) order: stable when that is element order. Neither the first pass (its
    // input is in text order) nor the partition of the ISA pairs (their order within a window is free) needs to be
    // stable, so there an item just takes the next slot of its digit in its wave's counter (one LDS atomic).
    constexpr bool UNSTABLE = FROM_T || !VALS;
    uint32_t rk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        bool valid = item(k) < cnt;
        uint32_t d = digit(key[k]);
        rk[k] = 0;
        if (UNSTABLE) {
            if (valid) rk[k] = atomicAdd(&s_wcnt[wave][d], 1u);
            continue;
        }
        uint64_t peers = match_digit8(d, __ballot(valid));
        if (valid) {
            uint32_t before = count_below(peers);
            uint32_t c = s_wcnt[wave][d];
            rk[k] = c + before;
            s_wcnt[wave][d] = c + (uint32_t)__popcll(peers);   // every peer writes the same value
        }
    }
    __syncthreads();

    // per digit: wave prefixes, tile total (the lookback over earlier tiles of the block follows in the scatter)
    uint32_t tot = 0;
    if (tid < 256) {
#pragma unroll
        for (int w2 = 0; w2 < W; w2++) { uint32_t c = s_wcnt[w2][tid]; s_wcnt[w2][tid] = tot; tot += c; }
        st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, tr.tb == 0 ? ST_INC : ST_AGG, tot));
    }
    uint32_t tstart = block_excl_sum<BLOCK, false>(tid < 256 ? tot : 0u, s_scratch, nullptr);   // s_scratch: once
    if (tid < 256) s_tstart[tid] = tstart;
    __syncthreads();

    // local sorted position of each item; keys leave first, then values, through one LDS buffer
    uint32_t lidx[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t d = digit(key[k]);
        lidx[k] = s_tstart[d] + s_wcnt[wave][d] + rk[k];
    }
    // A full tile (every tile but a block's last) takes the same steps without a check per item: straight-line code
    // lets the compiler give every store its own registers, where the checked form reused one set and waited for
    // each store to finish before the next. The keys go into LDS before the lookback, so that the LDS writes and the
    // lookback's round trips overlap; only the global stores need the digit bases it gives.
    auto scatter = [&](auto full_tile) {
        constexpr bool FULL = decltype(full_tile)::value;
#pragma unroll
        for (int k = 0; k < ITEMS; k++)
            if (FULL || item(k) < cnt) s_buf[lidx[k]] = key[k];
        if (tid < 256) {
            uint32_t excl = 0;
            if (tr.tb > 0) {
                excl = lookback_win<LB>(status + tid, 256, tile, epoch);
                st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_INC, excl + tot));
            }
            const uint32_t dbv = VALS ? db[tid] : woff + (uint32_t)min((uint64_t)tid << wsh, (uint64_t)wn);
            s_gbase[tid] = dbv + excl - tstart;
        }
        __syncthreads();
        // the sorted keys come out of LDS into registers first, so that the values can go into the same LDS while
        // the keys' global stores are issued
        uint32_t dg[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t j = tid + k * BLOCK;
            key[k] = FULL || j < cnt ? s_buf[j] : 0ull;
            dg[k] = digit(key[k]);
        }
        if (VALS) {
            __syncthreads();
            uint32_t *s_v = reinterpret_cast<uint32_t *>(s_buf);
#pragma unroll
            for (int k = 0; k < ITEMS; k++)
                if (FULL || item(k) < cnt) s_v[lidx[k]] = val[k];
        }
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t j = tid + k * BLOCK;
            if (FULL || j < cnt) kout[s_gbase[dg[k]] + j] = key[k];
        }
        if (!VALS) return;
        uint32_t *s_v = reinterpret_cast<uint32_t *>(s_buf);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t j = tid + k * BLOCK;
            if (FULL || j < cnt) vout[s_gbase[dg[k]] + j] = s_v[j];
        }
    };
    if (cnt == (uint32_t)TILE) scatter(std::true_type{});
    else scatter(std::false_type{});
}

// ------------------------------------------------------------------ regroup (group heads, ISA, group lists)

// Rows a-1 and a of sorted keys K share a group when their keys agree above bit kshift (and, with blk, lie in the same
// block). The group of a row starts at the first row of its run of equal keys; its head is that row's SA position
// (pos[row], or the row itself when pos == nullptr).
//
// Head of the group holding the first row of every regroup tile (tile t starts at row t * RG_TILE), by a binary
// search over the keys: they are sorted within a block (and over the whole range when blk == nullptr). One thread per
// tile, so the tiles of k_regroup need nothing from each other.
__global__ __launch_bounds__(256) void k_regroup_carry(const uint64_t *__restrict__ K, const uint32_t *__restrict__ pos,
                                                       uint32_t kshift, const uint16_t *__restrict__ blk,
                                                       const BlockInfo *__restrict__ blocks, uint32_t ntiles,
                                                       uint32_t *__restrict__ carry) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntiles) return;
    const uint32_t a0 = t * (uint32_t)RG_TILE;
    const uint64_t v = K[a0] >> kshift;
    const uint32_t base = blk ? blocks[blk[a0]].off : 0u;
    // the first row in [base, a0] whose key is >= v lies in [lo, hi], and K[hi] >= v. Galloping down from a0 first:
    // the group mostly starts at a0 or a few rows before it, so few of the binary search's dependent loads remain
    uint32_t lo = base, hi = a0;
    for (uint32_t step = 1; hi > base; step <<= 1) {
        const uint32_t probe = hi - base > step ? hi - step : base;
        if ((K[probe] >> kshift) < v) {
            lo = probe + 1;
            break;
        }
        hi = probe;
    }
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if ((K[mid] >> kshift) < v) lo = mid + 1; else hi = mid;
    }
    carry[t] = pos ? pos[lo] : lo;
}

// Sorted keys K[0..m) with values V (suffixes). Row a goes to SA position pos[a] (pos == nullptr: a). Writes
// SA[pos[a]] = V[a] (when SA != nullptr), ISA[V[a]] = head of its group (with pairs, the pair (V[a] << 32 | head)
// goes to pairs[a] instead, for k_isa_scatter after a partition by suffix), the BWT byte of row pos[a] into L (the
// key's low byte when T == nullptr, else T[V[a]-1]; the byte of a block's first suffix is never used), and appends
// every group of two or more rows as (head, size) to the small list (sh, ss; size <= SEG_GROUP) or the large list
// (lh, ls), one atomic per list and tile on cnt[0], cnt[1]; the lists' order is free. carry[tile]: the head of the
// group holding the tile's first row (k_regroup_carry).
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_regroup(const uint64_t *__restrict__ K, const uint32_t *__restrict__ V,
                                                   const uint32_t *__restrict__ pos, uint32_t m, uint32_t kshift,
                                                   const uint16_t *__restrict__ blk, uint32_t *__restrict__ SA,
                                                   uint32_t *__restrict__ ISA, uint64_t *__restrict__ pairs,
                                                   uint8_t *__restrict__ L, const uint8_t *__restrict__ T,
                                                   const uint32_t *__restrict__ carry, uint32_t *__restrict__ sh,
                                                   uint32_t *__restrict__ ss, uint32_t *__restrict__ lh,
                                                   uint32_t *__restrict__ ls, uint32_t *__restrict__ cnt_out) {
    // Rows come from HBM through LDS (coalesced loads in row order) and each thread works on ITEMS consecutive rows,
    // so one block-wide scan covers the tile. The pairs leave through LDS in row order. The LDS row of thread t
    // starts at t * (ITEMS + 1): the pad keeps a wave's lanes on distinct banks.
    constexpr int TILE = BLOCK * ITEMS, PITCH = ITEMS + 1;
    static_assert(ITEMS == 8, "a thread's BWT bytes leave as one 64-bit word");
    static_assert(TILE == RG_TILE, "k_regroup_carry's tiles");
    __shared__ uint64_t s_k[BLOCK * PITCH];      // keys, later the (suffix, head) pairs
    __shared__ uint32_t s_v[BLOCK * PITCH];      // suffixes
    __shared__ uint16_t s_b[BLOCK * PITCH];       // block ids (batches)
    __shared__ uint32_t s_scratch[32];
    __shared__ uint32_t s_imax[BLOCK];
    __shared__ uint32_t s_base[2];
    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const size_t t0 = (size_t)tile * TILE;
    const uint32_t cnt = (uint32_t)(t0 + TILE <= m ? TILE : m - t0);
    auto at = [](uint32_t j) { return (j / ITEMS) * PITCH + j % ITEMS; };
    {   // all of the tile's loads in flight before the first LDS write (a full tile loads without checks)
        uint64_t kk[ITEMS];
        uint32_t vv[ITEMS];
        uint16_t bb[ITEMS];
        auto load = [&](auto full_tile) {
            constexpr bool FULL = decltype(full_tile)::value;
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                const uint32_t j = k * BLOCK + tid;
                const size_t a = FULL || j < cnt ? t0 + j : t0;   // a row of the tile either way
                kk[k] = K[a];
                vv[k] = V[a];
                bb[k] = blk ? blk[a] : (uint16_t)0;
            }
        };
        if (cnt == (uint32_t)TILE) load(std::true_type{});
        else load(std::false_type{});
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t j = k * BLOCK + tid;
            if (j < cnt) {
                s_k[at(j)] = kk[k];
                s_v[at(j)] = vv[k];
                if (blk) s_b[at(j)] = bb[k];
            }
        }
    }
    // the rows on either side of the tile
    const bool has_prev = t0 > 0, has_next = t0 + cnt < m;
    const uint64_t kprev_t = has_prev ? K[t0 - 1] >> kshift : 0, knext_t = has_next ? K[t0 + cnt] >> kshift : 0;
    const uint16_t bprev_t = has_prev && blk ? blk[t0 - 1] : 0, bnext_t = has_next && blk ? blk[t0 + cnt] : 0;
    const uint32_t head_in = carry[tile];
    __syncthreads();

    const uint32_t r0 = (uint32_t)tid * ITEMS;   // this thread's first row in the tile
    uint64_t kc[ITEMS + 2];                      // [0]: row r0 - 1, [ITEMS + 1]: row r0 + ITEMS
    uint16_t bk[ITEMS + 2];
    uint32_t lp[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS + 2; k++) {
        const int32_t j = (int32_t)r0 + k - 1;
        if (j < 0) { kc[k] = kprev_t; bk[k] = bprev_t; }
        else if ((uint32_t)j >= cnt) { kc[k] = knext_t; bk[k] = bnext_t; }
        else { kc[k] = s_k[at((uint32_t)j)] >> kshift; bk[k] = blk ? s_b[at((uint32_t)j)] : 0; }
    }
    uint8_t lbyte[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t j = r0 + k;
        lbyte[k] = j < cnt ? (uint8_t)s_k[at(j)] : 0;
        const uint32_t pj = pos ? pos[t0 + (j < cnt ? j : cnt - 1)] : (uint32_t)(t0 + j);   // no branch on the load
        lp[k] = j < cnt ? pj : 0u;
    }
    uint32_t flags = 0, ends = 0;                // bit k: row k starts a group / ends a group of two or more
    uint32_t tmax = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t j = r0 + k;
        if (j >= cnt) continue;
        const bool pv = j > 0 || has_prev, nx = j + 1 < cnt || has_next;
        const bool flag = !pv || kc[k] != kc[k + 1] || bk[k] != bk[k + 1];
        const bool flagn = !nx || kc[k + 2] != kc[k + 1] || bk[k + 2] != bk[k + 1];
        if (flag) { flags |= 1u << k; tmax = lp[k]; }   // SA positions increase along the rows
        else if (flagn) ends |= 1u << k;
    }
    uint32_t ex_s, in_m, tot_s, tot_m;
    block_scan_sum_max<BLOCK>(0u, tmax, ex_s, in_m, tot_s, tot_m, s_scratch);
    s_imax[tid] = in_m;
    __syncthreads();
    // the head before this thread's first row: the tile's carry and the threads before this one
    uint32_t head = head_in;
    if (tid > 0 && s_imax[tid - 1] > head) head = s_imax[tid - 1];
    uint32_t eh[ITEMS], es[ITEMS];
    uint32_t ns = 0, nl = 0;
    uint64_t lbytes = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t j = r0 + k;
        eh[k] = 0;
        es[k] = 0;
        if (j >= cnt) continue;
        if (flags & (1u << k)) head = lp[k] > head ? lp[k] : head;
        const uint32_t i = s_v[at(j)];
        if (SA) SA[lp[k]] = i;
        if (pairs) s_k[at(j)] = ((uint64_t)i << 32) | head;
        else ISA[i] = head;
        const uint8_t c = T ? (i ? T[i - 1] : (uint8_t)0) : lbyte[k];
        if (pos) L[lp[k]] = c;
        else lbytes |= (uint64_t)c << (8 * k);
        if (ends & (1u << k)) {
            eh[k] = head;
            es[k] = lp[k] - head + 1;
            if (es[k] > (uint32_t)SEG_GROUP) nl++; else ns++;
        }
    }
    if (!pos) {   // rows in SA order: the thread's ITEMS bytes of L are consecutive
        if (r0 + ITEMS <= cnt && ((t0 + r0) & 7) == 0) *reinterpret_cast<uint64_t *>(L + t0 + r0) = lbytes;
        else
            for (int k = 0; k < ITEMS; k++)
                if (r0 + k < cnt) L[t0 + r0 + k] = (uint8_t)(lbytes >> (8 * k));
    }
    // the tile's groups into the lists: one atomic per list
    uint32_t tot;
    const uint32_t ex = block_excl_sum<BLOCK>(ns | (nl << 16), s_scratch, &tot);
    if (tid == 0) {
        s_base[0] = (tot & 0xFFFFu) ? atomicAdd(&cnt_out[0], tot & 0xFFFFu) : 0u;
        s_base[1] = (tot >> 16) ? atomicAdd(&cnt_out[1], tot >> 16) : 0u;
    }
    __syncthreads();
    uint32_t at_s = s_base[0] + (ex & 0xFFFFu), at_l = s_base[1] + (ex >> 16);
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        if (es[k] > (uint32_t)SEG_GROUP) { lh[at_l] = eh[k]; ls[at_l] = es[k]; at_l++; }
        else if (es[k]) { sh[at_s] = eh[k]; ss[at_s] = es[k]; at_s++; }
    }
    if (!pairs) return;
    // coalesced stores of the pairs in row order (the barrier above ordered the pair writes before these reads)
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t j = k * BLOCK + tid;
        if (j < cnt) pairs[t0 + j] = s_k[at(j)];
    }
}

// ------------------------------------------------------------------ ISA from partitioned pairs

// pairs: (suffix << 32 | head) partitioned by suffix window inside each block, so that every window occupies the
// positions of its own suffix range (the suffixes of a block are a permutation of its index range): the pairs at
// positions [p, q) carry suffixes of the windows that overlap [p, q). ISA[suffix] = head. A run of `super` tiles
// therefore writes a short stretch of ISA, and each run is written by one XCD (workgroups are dealt to the XCDs
// round-robin by id), whose L2 merges the writes into whole lines. A run covers a whole window of the largest block
// (at most 2 MB of ISA), so that no window's lines are written in part by several XCDs' L2s.
constexpr int SC_ITEMS = 16, SC_TILE = 256 * SC_ITEMS, NXCD = 8;
__global__ __launch_bounds__(256) void k_isa_scatter(const uint64_t *__restrict__ pairs, uint32_t N,
                                                     uint32_t *__restrict__ ISA, uint32_t super) {
    const uint32_t x = blockIdx.x % NXCD, k = blockIdx.x / NXCD;
    const size_t chunk = ((size_t)(k / super) * NXCD + x) * super + k % super;
    const size_t b0 = chunk * SC_TILE + threadIdx.x;
    uint64_t v[SC_ITEMS];
#pragma unroll
    for (int u = 0; u < SC_ITEMS; u++) {
        const size_t a = b0 + (size_t)u * 256;
        v[u] = a < N ? pairs[a] : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < SC_ITEMS; u++)
        if (v[u] != ~0ull) ISA[(uint32_t)(v[u] >> 32)] = (uint32_t)v[u];
}

// The same writes assembled in LDS: workgroup j owns the suffixes [lo, lo + ISA_SUB), lo = j * ISA_SUB, of the
// packed range. Its pairs all sit in the windows that overlap those suffixes (window d of block b holds the
// positions of its own suffix range), so it scans the positions of those windows, keeps the pairs whose suffix is
// its own, places each head in LDS and writes its stretch of ISA with 16-byte stores: every HBM line is written
// once and whole, with no partial-line writes for the L2 to merge. A window wider than ISA_SUB is scanned by each
// of the `super` workgroups that share it; they are dealt to one XCD one after the other (as in k_isa_scatter),
// so that its L2 can serve part of the re-reads.
constexpr int ISA_SUB = 32768, ISA_BLOCK = 1024, ISA_UNROLL = 4;
constexpr uint32_t ISA_LDS_SHARE = 4;   // most workgroups that scan one window
__global__ __launch_bounds__(ISA_BLOCK) void k_isa_lds(const uint64_t *__restrict__ pairs, uint32_t N,
                                                       const uint16_t *__restrict__ blk,
                                                       const BlockInfo *__restrict__ blocks,
                                                       uint32_t *__restrict__ ISA, uint32_t super) {
    __shared__ uint32_t s_isa[ISA_SUB];
    const uint32_t x = blockIdx.x % NXCD, k = blockIdx.x / NXCD;
    const size_t j = ((size_t)(k / super) * NXCD + x) * super + k % super;
    if (j * ISA_SUB >= N) return;
    const uint32_t lo = (uint32_t)j * ISA_SUB, cnt = N - lo < (uint32_t)ISA_SUB ? N - lo : (uint32_t)ISA_SUB;
    // positions [p0, p1): from the start of the window of suffix lo to the end of the window of suffix lo + cnt - 1
    const BlockInfo b0 = blocks[blk ? blk[lo] : 0u], b1 = blocks[blk ? blk[lo + cnt - 1] : 0u];
    const uint32_t s0 = window_shift(b0.n), s1 = window_shift(b1.n);
    const uint32_t p0 = (b0.off + ((lo - b0.off) >> s0 << s0)) & ~1u;   // even: two pairs per 16-byte load
    const uint64_t e1 = (uint64_t)b1.off + ((uint64_t)((lo + cnt - 1 - b1.off) >> s1) + 1) * (1ull << s1);
    const uint32_t p1 = (uint32_t)(e1 < (uint64_t)b1.off + b1.n ? e1 : (uint64_t)b1.off + b1.n);
    const int tid = threadIdx.x;
    // a pair outside [p0, p1) or of another window has its suffix outside [lo, lo + cnt) and is skipped
    for (uint32_t q0 = p0 + 2u * tid; q0 < p1; q0 += 2u * ISA_BLOCK * ISA_UNROLL) {
        uint64_t a[ISA_UNROLL], b[ISA_UNROLL];
#pragma unroll
        for (int u = 0; u < ISA_UNROLL; u++) {
            const uint32_t q = q0 + 2u * ISA_BLOCK * u;
            a[u] = b[u] = ~0ull;
            if (q + 1 < p1) {
                const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(pairs + q);
                a[u] = v.x;
                b[u] = v.y;
            } else if (q < p1) {
                a[u] = pairs[q];
            }
        }
#pragma unroll
        for (int u = 0; u < ISA_UNROLL; u++) {
            uint32_t r = (uint32_t)(a[u] >> 32) - lo;
            if (r < cnt) s_isa[r] = (uint32_t)a[u];
            r = (uint32_t)(b[u] >> 32) - lo;
            if (r < cnt) s_isa[r] = (uint32_t)b[u];
        }
    }
    __syncthreads();
    uint32_t *dst = ISA + lo;
    if (cnt == (uint32_t)ISA_SUB) {
        for (int i = tid; i < ISA_SUB / 4; i += ISA_BLOCK)
            reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(s_isa)[i];
    } else {
        for (uint32_t i = tid; i < cnt; i += ISA_BLOCK) dst[i] = s_isa[i];
    }
}

// ------------------------------------------------------------------ doubling round keys

// key = (group << 32) | sec (group: its index in the head-sorted list), sec = ISA[i+h] + h when i+h lies inside
// the block, else end-1-i (below every rank, shorter suffix first). orand[0] |= keys, orand[1] &= keys (which
// digits vary): one pair of atomics per workgroup.
__global__ __launch_bounds__(256) void k_round_keys(const uint32_t *__restrict__ pos, const uint32_t *__restrict__ grp,
                                                    uint32_t m, const uint32_t *__restrict__ SA,
                                                    const uint32_t *__restrict__ ISA, uint32_t n,
                                                    const uint16_t *__restrict__ blk,
                                                    const BlockInfo *__restrict__ blocks, uint32_t round,
                                                    uint64_t *__restrict__ K, uint32_t *__restrict__ V,
                                                    unsigned long long *__restrict__ orand) {
    __shared__ uint64_t s_o[4], s_a[4];
    uint64_t ko = 0, ka = ~0ull;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t a = (size_t)blockIdx.x * 256 + threadIdx.x; a < m; a += stride) {
        uint32_t p = pos[a];
        uint32_t i = SA[p];
        const BlockInfo bi = blocks[blk ? blk[p] : 0u];
        const uint32_t h = round_offset(bi, round), end = bi.off + bi.n;
        uint64_t j = (uint64_t)i + h;
        uint32_t sec = j < end ? ISA[j] + h : end - 1u - i;
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

// U of every block in libsais' convention from the row bytes L[p] = T[SA[p]-1]: U[0] = T[n-1], row p goes to U[p+1]
// before the primary row p0 (the rank of the block's first suffix) and to U[p] after it. 16 bytes per thread: when
// they lie in one block and U + o0 is 16-byte aligned they leave as one 16-byte store built from L[o0-1 .. o0+15],
// otherwise byte by byte.
__global__ __launch_bounds__(256) void k_bwt(const uint8_t *__restrict__ T, const uint8_t *__restrict__ L,
                                             const uint32_t *__restrict__ ISA, const uint16_t *__restrict__ blk,
                                             const BlockInfo *__restrict__ blocks, uint8_t *__restrict__ U,
                                             uint32_t N) {
    constexpr int CH = 64;   // output bytes per thread: four 16-byte vectors
    const size_t o0 = ((size_t)blockIdx.x * 256 + threadIdx.x) * CH;
    if (o0 >= N) return;
    auto info = [&](size_t o, uint32_t &off, uint32_t &n) {
        off = 0;
        n = N;
        if (blk) {
            const BlockInfo bi = blocks[blk[o]];
            off = bi.off;
            n = bi.n;
        }
    };
    const bool whole = o0 + CH <= N && (reinterpret_cast<uintptr_t>(U + o0) & 15u) == 0 &&
                       (!blk || blk[o0] == blk[o0 + CH - 1]);
    if (whole) {
        // byte lo of the block is L[lo - 1] (shifted) for 1 <= lo <= p0 and L[lo] after the primary row: each output
        // word is its own word of L, or that word shifted up one byte with the last byte of the word before it
        uint32_t off, n;
        info(o0, off, n);
        const uint32_t p0 = ISA[off] - off;
        uint32_t w[CH / 4 + 1];   // w[0]: the word before the chunk (L comes from hipMalloc: o0 is 16-aligned)
        w[0] = o0 >= 4 ? *reinterpret_cast<const uint32_t *>(L + o0 - 4) : 0u;
#pragma unroll
        for (int q = 0; q < CH / 16; q++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(L + o0 + 16 * q);
            w[1 + 4 * q] = v.x;
            w[2 + 4 * q] = v.y;
            w[3 + 4 * q] = v.z;
            w[4 + 4 * q] = v.w;
        }
        const uint8_t last = T[off + n - 1];
        uint32_t ow[CH / 4];
#pragma unroll
        for (int k = 0; k < CH / 4; k++) {
            const size_t lo = o0 + 4 * k - off;   // block position of the word's first byte
            const uint32_t sh = __builtin_amdgcn_alignbyte(w[k + 1], w[k], 3);   // L[lo - 1 .. lo + 2]
            uint32_t x;
            if (lo + 3 <= p0 && lo > 0) {
                x = sh;
            } else if (lo > p0) {
                x = w[k + 1];
            } else {
                x = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const size_t b = lo + j;
                    const uint32_t c = b == 0 ? last : (b <= p0 ? (sh >> (8 * j)) : (w[k + 1] >> (8 * j))) & 255u;
                    x |= (c & 255u) << (8 * j);
                }
            }
            ow[k] = x;
        }
#pragma unroll
        for (int q = 0; q < CH / 16; q++)
            *reinterpret_cast<uint4 *>(U + o0 + 16 * q) =
                make_uint4(ow[4 * q], ow[4 * q + 1], ow[4 * q + 2], ow[4 * q + 3]);
        return;
    }
    for (int i = 0; i < CH && o0 + i < N; i++) {
        const size_t o = o0 + i;
        uint32_t off, n;
        info(o, off, n);
        const uint32_t p0 = ISA[off] - off;
        const size_t lo = o - off;   // output byte lo of the block comes from row lo-1 (lo <= p0) or row lo (lo > p0)
        U[o] = lo == 0 ? T[off + n - 1] : (lo <= p0 ? L[o - 1] : L[o]);
    }
}

// I[g] for the g-th sampled index of the batch: block b with soff <= g, k = g - soff, I = 1 + rank of suffix k*r.
__global__ __launch_bounds__(256) void k_samples(const uint32_t *__restrict__ ISA, const BlockInfo *__restrict__ blocks,
                                                 uint32_t B, uint32_t count, int32_t *__restrict__ I) {
    size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= count) return;
    uint32_t lo = 0, hi = B;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (blocks[mid].soff <= g) lo = mid; else hi = mid;
    }
    const BlockInfo bi = blocks[lo];
    const size_t k = g - bi.soff;
    I[g] = (int32_t)(ISA[bi.off + k * bi.r] - bi.off + 1u);
}

// blk[o] = the block of packed index o: one workgroup per HIST_TILE indexes of a block (the text histograms' tiles).
__global__ __launch_bounds__(256) void k_fill_blk(uint16_t *__restrict__ blk, Tiling tl) {
    const uint32_t b = tile_block(tl.first, tl.B, blockIdx.x);
    const BlockInfo bi = tl.blocks[b];
    const size_t beg = (size_t)bi.off + (size_t)(blockIdx.x - tl.first[b]) * HIST_TILE;
    const size_t end = beg + HIST_TILE < (size_t)bi.off + bi.n ? beg + HIST_TILE : (size_t)bi.off + bi.n;
    for (size_t o = beg + threadIdx.x; o < end; o += 256) blk[o] = (uint16_t)b;
}

// ------------------------------------------------------------------ inverse BWT

// LF for every byte of every block's BWT: lf[q] = C_b[L[q]] + #{q' < q in the block : L[q'] = L[q]}, a global row.
// The row of L[q] in the block's full matrix is q before the primary index and q + 1 from it on.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_lf_pass(const uint8_t *__restrict__ L, Tiling tl,
                                                   const uint32_t *__restrict__ cbase, uint32_t *__restrict__ lf,
                                                   uint64_t *__restrict__ status, uint32_t *__restrict__ counter,
                                                   uint32_t epoch_arg, const uint32_t *__restrict__ depoch,
                                                   uint32_t ntiles) {
    const uint32_t epoch = depoch ? *depoch + epoch_arg : epoch_arg;   // a replayed graph reads its epoch here
    constexpr int W = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ uint32_t s_wcnt[W][256];
    __shared__ uint32_t s_base[256];
    __shared__ uint8_t s_l[TILE];          // the tile's bytes, staged with 16-byte loads
    __shared__ uint32_t s_tile;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    for (int i = tid; i < W * 256; i += BLOCK) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const TileRange tr = tile_range(tl, tile, TILE, 0);
    const uint32_t cnt = (uint32_t)(tr.end - tr.beg);
    stage_bytes<BLOCK>(L, tr.beg, cnt, tr.end, [&](uint32_t j, uint8_t c) { s_l[j] = c; });
    __syncthreads();
    uint32_t dg[ITEMS], rk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t local = wave * WAVE * ITEMS + k * WAVE + lane;
        bool valid = local < cnt;
        uint32_t d = valid ? s_l[local] : 0u;
        uint64_t peers = match_digit8(d, __ballot(valid));
        dg[k] = d;
        rk[k] = 0;
        if (valid) {
            uint32_t before = count_below(peers);
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
        if (tr.tb == 0) st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_INC, tot));
        else {
            st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_AGG, tot));
            excl = lookback_win<4>(status + tid, 256, tile, epoch);
            st_store(status + (size_t)tile * 256 + tid, st_pack(epoch, ST_INC, excl + tot));
        }
        s_base[tid] = cbase[(size_t)tr.b * 256 + tid] + excl;
    }
    __syncthreads();
    auto store = [&](auto full_tile) {   // a full tile stores without a check per item (no wait between the stores)
        constexpr bool FULL = decltype(full_tile)::value;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t local = wave * WAVE * ITEMS + k * WAVE + lane;
            if (FULL || local < cnt) lf[tr.beg + local] = s_base[dg[k]] + s_wcnt[wave][dg[k]] + rk[k];
        }
    };
    if (cnt == (uint32_t)TILE) store(std::true_type{});
    else store(std::false_type{});
}

// Byte of a row from its LF value (a global row): the c with C[c] <= v < C[c+1] (s_c holds the block's C[0..255]).
__device__ __forceinline__ uint8_t row_byte(const uint32_t *s_c, uint32_t v) {
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 128; step > 0; step >>= 1)
        if (s_c[lo + step] <= v) lo += step;
    return (uint8_t)lo;
}

// Two LF steps per load for the counting walk. For the stored row q of block b (row r = q - off, one more from P on):
// pair[q] = (stop << 40) | (L[q] << 32) | x, where L[q] is the byte of the step from r and LF(r) its landing row.
// stop: LF(r) is a splitter row (row % D == 0, or P), and then x = LF(r), where the walk ends; else x = LF(LF(r)).
// The second lookup follows LF, which is increasing over the rows of one byte, so the loads of a tile fall into one
// short run per byte value.
__global__ __launch_bounds__(LF_BLOCK) void k_lf_pairs(const uint32_t *__restrict__ lf, const uint8_t *__restrict__ L,
                                                       Tiling tl, uint32_t dshift, uint64_t *__restrict__ pair) {
    const uint32_t b = tile_block(tl.first, tl.B, blockIdx.x);
    const uint32_t tb = blockIdx.x - tl.first[b];
    const BlockInfo bi = tl.blocks[b];
    const uint32_t dmask = (1u << dshift) - 1u;
    const size_t beg = (size_t)bi.off + (size_t)tb * LF_TILE;
    const size_t end = (size_t)bi.off + bi.n;
    // every row's LF, then every second lookup and byte, then every pair: each phase's loads are in flight together
    // (the stored row of any landing row lies in the block, so a full tile gathers without checks)
    auto run = [&](auto full_tile) {
        constexpr bool FULL = decltype(full_tile)::value;
        uint32_t v[LF_ITEMS], x[LF_ITEMS];
        uint8_t c[LF_ITEMS];
#pragma unroll
        for (int k = 0; k < LF_ITEMS; k++) {
            const size_t q = beg + (size_t)k * LF_BLOCK + threadIdx.x;
            v[k] = FULL || q < end ? lf[q] : 0u;
        }
#pragma unroll
        for (int k = 0; k < LF_ITEMS; k++) {
            const size_t q = beg + (size_t)k * LF_BLOCK + threadIdx.x;
            const uint32_t row = v[k] - bi.roff;
            x[k] = 0;
            c[k] = 0;
            if (FULL || q < end) {
                x[k] = lf[bi.off + row - (row < bi.P ? 0u : 1u)];
                c[k] = L[q];
            }
        }
#pragma unroll
        for (int k = 0; k < LF_ITEMS; k++) {
            const size_t q = beg + (size_t)k * LF_BLOCK + threadIdx.x;
            if (FULL || q < end) {
                const uint32_t row = v[k] - bi.roff;
                const bool stop = (row & dmask) == 0 || row == bi.P;
                pair[q] = ((uint64_t)stop << 40) | ((uint64_t)c[k] << 32) | (stop ? v[k] : x[k]);
            }
        }
    };
    if (beg + LF_TILE <= end) run(std::true_type{});
    else run(std::false_type{});
}

// Splitters of block b: local j < nreg is the row j*D, j == nreg the row P (only when P % D != 0); global id
// spoff + j. From each, walk LF to the next splitter row of the block; link = (steps << 32) | next splitter, len =
// steps. The bytes met on the way (the text backwards from the splitter's suffix) are staged for the first cap steps
// (cap a multiple of 16) in 16-byte chunks, chunk-major: steps 16c..16c+15 of splitter id at stage[c * 16 * sstride
// + 16 * id]. The lanes of a wave walk consecutive splitters in step, so they store a chunk row together, 16 bytes
// each; cont keeps the local row of step cap of a longer sublist. The marker row P ends the list: link = (0, S).
// Workgroups take WALK_TILE splitters of one block.
//
// The walk is a chain of dependent loads, so its time is the longest sublist times the load latency. It moves two
// steps per load through the pair table (one when the first step lands on a splitter), and the next load is issued
// before the current bytes are looked up and stored.
__global__ __launch_bounds__(WALK_TILE) void k_walk_count(const uint64_t *__restrict__ pair, Tiling tl, uint32_t dshift,
                                                          uint32_t S, const uint32_t *__restrict__ cbase,
                                                          uint64_t *__restrict__ link, uint32_t *__restrict__ len,
                                                          uint8_t *__restrict__ stage, size_t sstride, uint32_t cap,
                                                          uint32_t *__restrict__ cont) {
    __shared__ uint32_t s_c[256];
    const uint32_t b = tile_block(tl.first, tl.B, blockIdx.x);
    const uint32_t tb = blockIdx.x - tl.first[b];
    const BlockInfo bi = tl.blocks[b];
    s_c[threadIdx.x] = cbase[(size_t)b * 256 + threadIdx.x];
    __syncthreads();
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) link[S] = S;
    const uint32_t dmask = (1u << dshift) - 1u;
    const uint32_t nsp = bi.nreg + ((bi.P & dmask) != 0 ? 1u : 0u);
    const uint32_t j = tb * WALK_TILE + threadIdx.x;
    if (j >= nsp) return;
    const uint32_t id = bi.spoff + j;
    const uint32_t sidP = bi.spoff + ((bi.P & dmask) == 0 ? (bi.P >> dshift) : bi.nreg);
    const uint32_t r0 = j < bi.nreg ? (j << dshift) : bi.P;
    if (r0 == bi.P) { link[id] = S; len[id] = 0; return; }
    uint32_t steps = 0, r = r0;
    uint64_t e = pair[bi.off + r0 - (r0 < bi.P ? 0u : 1u)];
    uint8_t *st = stage + (size_t)id * 16;
    const size_t cstride = sstride * 16;
    uint64_t lo = 0, hi = 0;                                   // the chunk being filled: steps (steps & ~15) + 0..15
    auto put = [&](uint32_t t, uint64_t bytes) {               // bytes of step t (t even) and on, within the chunk
        const uint64_t v = bytes << (8 * (t & 7));
        if (t & 8) hi |= v; else lo |= v;
    };
    auto store = [&](uint32_t c) {
        *reinterpret_cast<uint4 *>(st + (size_t)c * cstride) =
            make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
        lo = 0;
        hi = 0;
    };
    for (;;) {
        const uint32_t x = (uint32_t)e;
        const uint8_t b1 = (uint8_t)(e >> 32);
        if ((e >> 40) & 1u) {                                  // the first step lands on a splitter
            if (steps < cap) put(steps, b1);
            steps++;
            r = x - bi.roff;
            break;
        }
        const uint32_t rn = x - bi.roff;                       // the row after two steps
        const bool more = (rn & dmask) != 0 && rn != bi.P && steps + 2 <= bi.n;
        uint64_t en = 0;
        if (more) {
            if (steps + 2 == cap) cont[id] = rn;
            en = pair[bi.off + rn - (rn < bi.P ? 0u : 1u)];
        }
        if (steps < cap) {                                     // cap is a multiple of 16: both steps fit
            put(steps, (uint64_t)b1 | ((uint64_t)row_byte(s_c, x) << 8));
            if ((steps & 15) == 14) store(steps >> 4);
        }
        steps += 2;
        r = rn;
        if (!more) break;
        e = en;
    }
    if ((steps & 15) != 0 && steps <= cap) store((steps - 1) >> 4);   // the last, partial chunk
    const uint32_t nxt = r == bi.P ? sidP : bi.spoff + (r >> dshift);
    link[id] = ((uint64_t)steps << 32) | nxt;
    len[id] = steps;
}

// m-hop pointer jumping: out[s] = the node HOPS links on from s with the lengths on the way summed (the end S stops
// a walk). Links that each cover d list nodes cover HOPS * d after one step.
constexpr int RANK_HOPS = 4;
__global__ __launch_bounds__(256) void k_rank_jump(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                   uint32_t S) {
    size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s > S) return;
    uint64_t v = in[s];
    uint32_t ptr = (uint32_t)v;
    uint64_t acc = v >> 32;
#pragma unroll
    for (int hop = 1; hop < RANK_HOPS; hop++) {
        if (ptr == S) break;
        const uint64_t w = in[ptr];
        acc += w >> 32;
        ptr = (uint32_t)w;
    }
    out[s] = (acc << 32) | ptr;
}

// A block's BWT is valid when the list from its row 0 (splitter spoff) reaches the end with n steps.
__global__ __launch_bounds__(256) void k_check_lists(const uint64_t *__restrict__ link,
                                                     const BlockInfo *__restrict__ blocks, uint32_t B, uint32_t S,
                                                     uint32_t *__restrict__ bad) {
    size_t b = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const uint64_t l = link[blocks[b].spoff];
    bad[b] = ((uint32_t)l == S && (uint32_t)(l >> 32) == blocks[b].n) ? 0u : 1u;
}

// Write every sublist: its splitter's suffix starts at block position pos = link >> 32 (after ranking), and step t of
// the walk is the byte at pos-1-t. The staged steps go through LDS, EMIT_ROWS steps of the workgroup's WALK_TILE
// sublists at a time: each lane reads its sublist's 16-byte chunks of those steps (only chunks it staged), and each
// wave then writes the bytes of one sublist at a time as one contiguous run. A sublist longer than cap walks on from
// cont.
constexpr int EMIT_ROWS = 64, EMIT_PITCH = EMIT_ROWS + 16;   // a sublist's 64 steps per LDS row; the pitch keeps
                                                            // the lanes' 16-byte writes free of bank conflicts
__global__ __launch_bounds__(WALK_TILE) void k_walk_emit(const uint32_t *__restrict__ lf, Tiling tl, uint32_t dshift,
                                                         uint32_t S, const uint64_t *__restrict__ link,
                                                         const uint32_t *__restrict__ len,
                                                         const uint32_t *__restrict__ cbase,
                                                         const uint8_t *__restrict__ stage, size_t sstride,
                                                         uint32_t cap, const uint32_t *__restrict__ cont,
                                                         const uint32_t *__restrict__ bad, uint8_t *__restrict__ out) {
    __shared__ uint32_t s_c[256];
    __shared__ __align__(16) uint8_t s_rows[WALK_TILE * EMIT_PITCH];   // [sublist][step of the window]
    __shared__ uint32_t s_max;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const uint32_t b = tile_block(tl.first, tl.B, blockIdx.x);
    const uint32_t tb = blockIdx.x - tl.first[b];
    const BlockInfo bi = tl.blocks[b];
    if (bad[b]) {   // the string is no BWT: it decodes to zeros (the caller's checksum rejects them)
        // the block's tiles cover its positions WALK_TILE * D at a time (nsp * D > n); uniform over the workgroup
        const size_t z0 = ((size_t)tb * WALK_TILE) << dshift;
        const size_t z1 = z0 + ((size_t)WALK_TILE << dshift) < bi.n ? z0 + ((size_t)WALK_TILE << dshift) : bi.n;
        for (size_t o = z0 + tid; o < z1; o += WALK_TILE) out[bi.off + o] = 0;
        return;
    }
    s_c[tid] = cbase[(size_t)b * 256 + tid];
    if (tid == 0) s_max = 0;
    const uint32_t dmask = (1u << dshift) - 1u;
    const uint32_t nsp = bi.nreg + ((bi.P & dmask) != 0 ? 1u : 0u);
    const uint32_t j0 = tb * WALK_TILE;
    const uint32_t cnt = nsp - j0 < (uint32_t)WALK_TILE ? nsp - j0 : (uint32_t)WALK_TILE;
    const uint32_t id0 = bi.spoff + j0;
    uint32_t steps = 0, pos = 0;
    if ((uint32_t)tid < cnt) {
        steps = len[id0 + tid];
        pos = (uint32_t)(link[id0 + tid] >> 32);
        if (pos > bi.n || steps > pos) steps = 0;           // only a string that is no BWT gets here
    }
    const uint32_t staged = steps < cap ? steps : cap;
    __syncthreads();
    if (staged) atomicMax(&s_max, staged);
    __syncthreads();
    const uint32_t rows = s_max;
    for (uint32_t t0 = 0; t0 < rows; t0 += EMIT_ROWS) {
        const uint32_t nr = rows - t0 < (uint32_t)EMIT_ROWS ? rows - t0 : (uint32_t)EMIT_ROWS;
        if ((uint32_t)tid < cnt) {
            // chunks t0/16 .. t0/16 + EMIT_ROWS/16 - 1 of this lane's sublist, all loads in flight together
            const uint8_t *src = stage + (size_t)(t0 >> 4) * 16 * sstride + (size_t)(id0 + tid) * 16;
            uint4 v[EMIT_ROWS / 16];
#pragma unroll
            for (int q = 0; q < EMIT_ROWS / 16; q++)
                v[q] = t0 + 16 * q < staged ? *reinterpret_cast<const uint4 *>(src + (size_t)q * 16 * sstride)
                                            : make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < EMIT_ROWS / 16; q++) *reinterpret_cast<uint4 *>(s_rows + tid * EMIT_PITCH + 16 * q) = v[q];
        }
        __syncthreads();
        // sublist k of the wave is lane k's: its staged steps and position come from that lane's registers as scalars
        uint8_t *const ob = out + bi.off - 1;
        const uint32_t t = t0 + (uint32_t)lane;
        const bool row_ok = (uint32_t)lane < nr;
        // only the wave's sublists with steps in this window
        for (unsigned long long act = __ballot(staged > t0); act; act &= act - 1) {
            const int k = __builtin_ctzll(act);
            const uint32_t sk = __builtin_amdgcn_readlane(staged, k), pk = __builtin_amdgcn_readlane(pos, k);
            if (t < sk && row_ok) ob[pk - t] = s_rows[(wave * WAVE + k) * EMIT_PITCH + lane];
        }
        __syncthreads();
    }
    if (steps > cap) {
        uint8_t *o = out + bi.off + pos - 1;
        uint32_t r = cont[id0 + tid];
        for (uint32_t t = cap; t < steps; t++) {
            const uint32_t v = lf[bi.off + r - (r < bi.P ? 0u : 1u)];
            o[-(int64_t)t] = row_byte(s_c, v);
            r = v - bi.roff;
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

// The text histograms run four workgroups per CU (their LDS allows it): each one's LDS atomics wait long, and more
// waves per SIMD hide that (k_hist_T: 0.131 -> 0.103 ms on a 32 MiB block against 512 workgroups).
constexpr uint32_t TEXT_HIST_GRID = 1024;
void launch_byte_hist(const uint8_t *T, Tiling tl, uint32_t ntiles, uint32_t *counts, hipStream_t s) {
    // Each workgroup adds its bins to its blocks' 256 counters: more than TEXT_HIST_GRID workgroups on one block
    // queue on those atomics (0.030 -> 0.046 ms on a 32 MiB block at 2048), while a batch spreads them over its
    // blocks' counters and gains from more (1.66 -> 1.11 ms per step on 512 blocks at 4096).
    const uint32_t want = TEXT_HIST_GRID * (tl.B < 4u ? (tl.B ? tl.B : 1u) : 4u);
    const uint32_t grid = ntiles < want ? ntiles : want;
    if (grid) k_byte_hist<<<grid, 256, 0, s>>>(T, tl, ntiles, counts);
}
void launch_scan_digits(uint32_t *counts, uint32_t nsets, uint32_t sets_per_block, const BlockInfo *blocks,
                        int base_mode, uint32_t *tot, uint32_t *excl, hipStream_t s) {
    if (nsets) k_scan_digits<<<nsets, 256, 0, s>>>(counts, sets_per_block, blocks, base_mode, tot, excl);
}
void launch_hist_T(const uint8_t *T, Tiling tl, uint32_t ntiles, const uint8_t *codes, uint32_t P, uint32_t *counts,
                   hipStream_t s) {
    const uint32_t grid = ntiles < TEXT_HIST_GRID ? ntiles : TEXT_HIST_GRID;
    if (grid) k_hist_T<<<grid, 256, 0, s>>>(T, tl, ntiles, codes, P, counts);
}
void launch_hist_gram(const uint8_t *T, Tiling units, uint32_t nunits, const uint8_t *codes, const uint32_t *tot,
                      uint32_t P, uint32_t *counts, hipStream_t s) {
    if (nunits) k_hist_gram<<<nunits, GRAM_BLOCK, 0, s>>>(T, units, codes, tot, P, counts);
}
void launch_hist_keys(const uint64_t *K, uint32_t m, const Shifts &sh, uint32_t *counts, uint32_t grid,
                      hipStream_t s) {
    k_hist_keys<<<grid, 256, 0, s>>>(K, m, sh, counts);
}

// The radix pass configurations the tuner chooses from: (workgroup size, items per thread, lookback window).
const SortCfg SORT_CFGS[SORT_NCFG] = {{512, 16, 4}, {256, 16, 4}, {256, 16, 1}, {512, 8, 4},
                                      {256, 12, 4}, {256, 8, 4}, {512, 16, 1}, {1024, 12, 4},
                                      {1024, 16, 4}, {1024, 16, 2}, {1024, 16, 8}};
uint32_t sort_tile(int cfg) { return (uint32_t)(SORT_CFGS[cfg].block * SORT_CFGS[cfg].items); }

template <int BLOCK, int ITEMS, int LB>
static void sort_pass_cfg(const SortArgs &a, uint32_t ntiles, hipStream_t s) {
    if (a.from_t)
        k_sort_pass<BLOCK, ITEMS, LB, true><<<ntiles, BLOCK, 0, s>>>(a.kin, a.vin, a.kout, a.vout, a.m, a.shift, a.dbase,
                                                                    a.dstride, a.tiling, a.status, a.counter, a.epoch,
                                                                    a.depoch, ntiles, a.T, a.codes);
    else
        k_sort_pass<BLOCK, ITEMS, LB, false><<<ntiles, BLOCK, 0, s>>>(a.kin, a.vin, a.kout, a.vout, a.m, a.shift,
                                                                     a.dbase, a.dstride, a.tiling, a.status,
                                                                     a.counter, a.epoch, a.depoch, ntiles, a.T,
                                                                     a.codes);
}
void launch_sort_pass(const SortArgs &a, uint32_t ntiles, hipStream_t s) {
    if (!ntiles) return;
    switch (a.cfg) {
    case 1: sort_pass_cfg<256, 16, 4>(a, ntiles, s); break;
    case 2: sort_pass_cfg<256, 16, 1>(a, ntiles, s); break;
    case 3: sort_pass_cfg<512, 8, 4>(a, ntiles, s); break;
    case 4: sort_pass_cfg<256, 12, 4>(a, ntiles, s); break;
    case 5: sort_pass_cfg<256, 8, 4>(a, ntiles, s); break;
    case 6: sort_pass_cfg<512, 16, 1>(a, ntiles, s); break;
    case 7: sort_pass_cfg<1024, 12, 4>(a, ntiles, s); break;
    case 8: sort_pass_cfg<1024, 16, 4>(a, ntiles, s); break;
    case 9: sort_pass_cfg<1024, 16, 2>(a, ntiles, s); break;
    case 10: sort_pass_cfg<1024, 16, 8>(a, ntiles, s); break;
    default: sort_pass_cfg<512, 16, 4>(a, ntiles, s); break;
    }
}
void launch_partition_pairs(const uint64_t *in, uint64_t *out, uint32_t N, const Tiling &tl, uint32_t ntiles,
                            uint64_t *status, uint32_t *counter, uint32_t epoch, const uint32_t *depoch,
                            hipStream_t s) {
    if (!ntiles) return;
    k_sort_pass<PART_BLOCK, PART_ITEMS, 4, false, false><<<ntiles, PART_BLOCK, 0, s>>>(
        in, nullptr, out, nullptr, N, 0, nullptr, 0, tl, status, counter, epoch, depoch, ntiles, nullptr, nullptr);
}
void launch_isa_scatter(const uint64_t *pairs, uint32_t N, uint32_t maxn, const uint16_t *blk,
                        const BlockInfo *blocks, uint32_t *ISA, hipStream_t s) {
    if (!N) return;
    // the largest block's window (256 windows per block, as the ISA partition cuts them)
    const uint32_t ws = maxn > 256 ? 24u - (uint32_t)__builtin_clz(maxn - 1u) : 0u;
    // Windows of up to four workgroups' suffixes are assembled in LDS. Each of a window's workgroups reads all of
    // its pairs, and the L2 serves few of the re-reads (the 32 workgroups of an XCD stream eight or more windows at
    // once). Up to four per window it still gains (0.198 -> 0.186 ms on a 32 MiB block, four per window; 7.26 ->
    // 3.63 ms per step on 512 blocks of up to 16 MiB, two at most); at sixteen the re-reads cost more than the
    // partial-line writes (1.31 -> 2.40 ms on a 128 MiB block). Wider windows keep the scatter below.
    if ((1u << ws) <= ISA_LDS_SHARE * (uint32_t)ISA_SUB) {
        // workgroups per window (1 when a window is no wider than a workgroup's suffixes)
        const uint32_t super = (1u << ws) > (uint32_t)ISA_SUB ? (1u << ws) / ISA_SUB : 1u;
        const size_t subs = ((size_t)N + ISA_SUB - 1) / ISA_SUB;
        const size_t runs = (subs + super - 1) / super;
        const size_t grid = (runs + NXCD - 1) / NXCD * NXCD * super;
        k_isa_lds<<<(uint32_t)grid, ISA_BLOCK, 0, s>>>(pairs, N, blk, blocks, ISA, super);
        return;
    }
    // runs of 32 to 128 tiles
    const uint32_t wt = (1u << ws) / SC_TILE;
    const uint32_t super = wt < 32u ? 32u : (wt > 128u ? 128u : wt);
    const size_t chunks = ((size_t)N + SC_TILE - 1) / SC_TILE;
    const size_t runs = (chunks + super - 1) / super;
    const size_t grid = (runs + NXCD - 1) / NXCD * NXCD * super;
    k_isa_scatter<<<(uint32_t)grid, 256, 0, s>>>(pairs, N, ISA, super);
}
void launch_regroup(const RegroupArgs &a, hipStream_t s) {
    const uint32_t ntiles = (a.m + RG_TILE - 1) / RG_TILE;
    if (!ntiles) return;
    k_regroup_carry<<<(ntiles + 255) / 256, 256, 0, s>>>(a.K, a.pos, a.kshift, a.blk, a.blocks, ntiles, a.carry);
    k_regroup<RG_BLOCK, RG_ITEMS><<<ntiles, RG_BLOCK, 0, s>>>(a.K, a.V, a.pos, a.m, a.kshift, a.blk, a.SA, a.ISA,
                                                             a.pairs, a.L, a.T, a.carry, a.sh, a.ss, a.lh, a.ls,
                                                             a.cnt);
}
void launch_round_keys(const uint32_t *pos, const uint32_t *grp, uint32_t m, const uint32_t *SA, const uint32_t *ISA,
                       uint32_t n, const uint16_t *blk, const BlockInfo *blocks, uint32_t round, uint64_t *K, uint32_t *V,
                       unsigned long long *orand, hipStream_t s) {
    uint32_t grid = (m + 255) / 256;
    if (grid > 2048) grid = 2048;
    if (grid) k_round_keys<<<grid, 256, 0, s>>>(pos, grp, m, SA, ISA, n, blk, blocks, round, K, V, orand);
}
void launch_bwt(const uint8_t *T, const uint8_t *L, const uint32_t *ISA, const uint16_t *blk, const BlockInfo *blocks,
                uint8_t *U, uint32_t N, hipStream_t s) {
    if (N) k_bwt<<<(uint32_t)(((size_t)N + 256 * 64 - 1) / (256 * 64)), 256, 0, s>>>(T, L, ISA, blk, blocks, U, N);
}
void launch_samples(const uint32_t *ISA, const BlockInfo *blocks, uint32_t B, uint32_t count, int32_t *I,
                    hipStream_t s) {
    if (count) k_samples<<<(count + 255) / 256, 256, 0, s>>>(ISA, blocks, B, count, I);
}
void launch_fill_blk(uint16_t *blk, Tiling tl, uint32_t ntiles, hipStream_t s) {
    if (ntiles) k_fill_blk<<<ntiles, 256, 0, s>>>(blk, tl);
}
void launch_lf_pass(const uint8_t *L, Tiling tl, uint32_t ntiles, const uint32_t *cbase, uint32_t *lf,
                    uint64_t *status, uint32_t *counter, uint32_t epoch, const uint32_t *depoch, hipStream_t s) {
    if (ntiles)
        k_lf_pass<LF_BLOCK, LF_ITEMS><<<ntiles, LF_BLOCK, 0, s>>>(L, tl, cbase, lf, status, counter, epoch, depoch,
                                                                 ntiles);
}
// Device-side epochs of a replayed graph: *depoch moves on by EPOCH_STRIDE per replay, within the upper half of
// the epoch space (host epochs stay in the lower half).
__global__ void k_epoch_bump(uint32_t *depoch) {
    const uint32_t v = *depoch + EPOCH_STRIDE;
    *depoch = (v & (EPOCH_HALF - 1)) | EPOCH_HALF;
}
void launch_epoch_bump(uint32_t *depoch, hipStream_t s) { k_epoch_bump<<<1, 1, 0, s>>>(depoch); }
void launch_lf_pairs(const uint32_t *lf, const uint8_t *L, Tiling tl, uint32_t ntiles, uint32_t dshift,
                     uint64_t *pair, hipStream_t s) {
    if (ntiles) k_lf_pairs<<<ntiles, LF_BLOCK, 0, s>>>(lf, L, tl, dshift, pair);
}
void launch_walk_count(const uint64_t *pair, Tiling tl, uint32_t ntiles, uint32_t dshift, uint32_t S,
                       const uint32_t *cbase, uint64_t *link, uint32_t *len, uint8_t *stage, size_t sstride,
                       uint32_t cap, uint32_t *cont, hipStream_t s) {
    if (ntiles)
        k_walk_count<<<ntiles, WALK_TILE, 0, s>>>(pair, tl, dshift, S, cbase, link, len, stage, sstride, cap, cont);
}
uint32_t rank_jump_iters(uint32_t S) {
    // reach after k steps: RANK_HOPS^k list nodes; a list has at most S nodes
    uint32_t k = 0;
    uint64_t reach = 1;
    while (reach < (uint64_t)S + 1) { reach *= RANK_HOPS; k++; }
    return k + 1;
}
void launch_rank_jump(const uint64_t *in, uint64_t *out, uint32_t S, hipStream_t s) {
    k_rank_jump<<<(S + 1 + 255) / 256, 256, 0, s>>>(in, out, S);
}
void launch_check_lists(const uint64_t *link, const BlockInfo *blocks, uint32_t B, uint32_t S, uint32_t *bad,
                        hipStream_t s) {
    if (B) k_check_lists<<<(B + 255) / 256, 256, 0, s>>>(link, blocks, B, S, bad);
}
void launch_walk_emit(const uint32_t *lf, Tiling tl, uint32_t ntiles, uint32_t dshift, uint32_t S,
                      const uint64_t *link, const uint32_t *len, const uint32_t *cbase, const uint8_t *stage,
                      size_t sstride, uint32_t cap, const uint32_t *cont, const uint32_t *bad, uint8_t *out,
                      hipStream_t s) {
    if (ntiles)
        k_walk_emit<<<ntiles, WALK_TILE, 0, s>>>(lf, tl, dshift, S, link, len, cbase, stage, sstride, cap, cont, bad,
                                                 out);
}
void launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s) {
    size_t n16 = bytes / 16;
    if (n16) k_copy16<<<2048, 256, 0, s>>>((uint4 *)dst, (const uint4 *)src, n16);
}
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (n) k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
}

}  // namespace kbwt
