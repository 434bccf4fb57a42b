// Doubling rounds over groups (gfx950).
//
// After the first sort, the suffixes that still share their first h symbols form groups: contiguous suffix-array
// ranges [head, head + size) with size >= 2, listed as (head, size) descriptors. A round with offset h gives every
// suffix i of a group the key sec(i) = ISA[i + h] + h (or end - 1 - i when i + h lies past its block's end, below every
// rank, shorter suffix first), sorts each group by sec, and splits it where sec changes: the subgroup heads become
// the new ranks (ISA), the rows are rewritten in SA with their BWT bytes, and subgroups of two or more form the next
// round's list. All keys of a round are gathered before any rank of the round changes.
//
// Small groups (at most SEG_GROUP suffixes) never leave one workgroup: a tile owns the groups whose first element
// offset (prefix sum of sizes over the list) lies in its SEG_TILE-wide range, gathers them into LDS and sorts them
// there by (group, sec) with an LDS radix sort or direct ranking, so a round costs a few bytes per suffix instead of
// the global radix passes. Larger groups take the global path (k_round_keys, radix passes, k_regroup), whose
// regroup appends the subgroups to the next round's lists itself.
#include "kbwt_device.h"
#include "kbwt_internal.h"

namespace kbwt {

// First index g in [0, G) with go[g] >= x (G when none).
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *go, uint32_t G, uint32_t x) {
    uint32_t lo = 0, hi = G;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (go[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Exclusive prefix sum of the group sizes: go[g] = sizes before g; the last tile writes the total to *tot. Also the
// first group of every SEG_TILE-row tile of k_seg_round, tstart[t] = the first g with go[g] >= t * SEG_TILE
// (tstart[tiles] = G): group g, rows [go[g], go[g] + size) with size <= SEG_TILE, holds at most one tile start t *
// SEG_TILE, whose first group is g when the tile starts with it and g + 1 otherwise.
__global__ __launch_bounds__(256) void k_scan_groups(const uint32_t *__restrict__ gs, uint32_t G,
                                                     uint32_t *__restrict__ go, uint32_t *__restrict__ tot,
                                                     uint32_t *__restrict__ tstart,
                                                     uint64_t *__restrict__ status, uint32_t *__restrict__ counter,
                                                     uint32_t epoch, uint32_t ntiles) {
    constexpr int ITEMS = 8, TILE = 256 * ITEMS;
    __shared__ uint32_t s_scratch[8];
    __shared__ uint32_t s_tile, s_ex;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = next_tile(counter, ntiles);
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * TILE + (size_t)tid * ITEMS;   // blocked: thread owns ITEMS consecutive
    uint32_t v[ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        v[k] = base + k < G ? gs[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t ex = block_excl_sum<256>(sum, s_scratch, &total);
    if (tid < WAVE) {
        uint32_t prev = 0;
        if (tile == 0) {
            if (tid == 0) st_store(status, st_pack(epoch, ST_INC, total));
        } else {
            if (tid == 0) st_store(status + tile, st_pack(epoch, ST_AGG, total));
            prev = lookback_wave<false>(status, tile, epoch);
            if (tid == 0) st_store(status + tile, st_pack(epoch, ST_INC, prev + total));
        }
        if (tid == 0) {
            s_ex = prev;
            if (tile == ntiles - 1) *tot = prev + total;
        }
    }
    __syncthreads();
    uint32_t run = s_ex + ex;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const size_t g = base + k;
        if (g < G) {
            go[g] = run;
            const uint32_t t = (run + (uint32_t)SEG_TILE - 1) / (uint32_t)SEG_TILE;   // first tile start >= run
            if (tstart) {
                if ((uint64_t)t * SEG_TILE < (uint64_t)run + v[k])
                    tstart[t] = (uint64_t)t * SEG_TILE == run ? (uint32_t)g : (uint32_t)g + 1;
                if (g + 1 == G) tstart[(run + v[k] + (uint32_t)SEG_TILE - 1) / (uint32_t)SEG_TILE] = G;
            }
        }
        run += v[k];
    }
}

// Rows of the large groups as element lists for the global path: pos = head + k, grp = the group's index in the
// list (sorted by head), in list order.
// lo[g] = first row of large group g in the element list (host prefix sums), G large groups, m rows.
__global__ __launch_bounds__(256) void k_large_elems(const uint32_t *__restrict__ lh, const uint32_t *__restrict__ lo,
                                                     uint32_t G, uint32_t m, uint32_t *__restrict__ pos,
                                                     uint32_t *__restrict__ grp) {
    for (size_t a = (size_t)blockIdx.x * 256 + threadIdx.x; a < m; a += (size_t)gridDim.x * 256) {
        uint32_t g = lower_bound_u32(lo, G, (uint32_t)a + 1) - 1;   // last g with lo[g] <= a
        pos[a] = lh[g] + ((uint32_t)a - lo[g]);
        grp[a] = g;   // the list is in head order, so the index orders groups as their heads do, in fewer bits
    }
}

// One tile of small groups. MODE 0 gathers every row's key and records in tmax the rows of the tile's largest group
// that splits (whose keys are not all equal; 0 when none does). The split sorts each group by its key in LDS,
// rewrites SA, ISA and the BWT bytes of its rows, and writes the subgroups of two or more into the tile's own area of
// the emission buffers (nh, ns at tile * SEG_TILE; ncnt[tile] counts them), which k_compact_emits packs into the next
// list (one shared counter for all tiles would serialise every workgroup on a single atomic). MODE 2 splits the
// tiles whose splitting groups have at most rank_max rows (empty tiles and tiles where nothing splits included) by
// direct ranking; its small LDS footprint lets eight workgroups share a CU. MODE 1 splits the other tiles by an LDS
// radix sort.
//
// The BWT byte of a row, T[i - 1] for suffix i, is written to L when the row ends up alone (its SA position is final
// then). In the first round (CARRY) it needs no read of T: L holds the byte of every row after the first sort, the
// gather reads it along with SA (lrow, in list order), and every row that moves takes it along (L[r] stays the byte
// before the suffix at SA row r). Later rounds move rows many times before they end up alone, so there a row reads T
// once, at the end, instead.
template <int MODE, bool CARRY>
__global__ __launch_bounds__(256) void k_seg_round(const uint32_t *__restrict__ gh, const uint32_t *__restrict__ go,
                                                   const uint32_t *__restrict__ tstart, uint32_t G, uint32_t E, uint32_t *__restrict__ SA,
                                                   uint32_t *__restrict__ ISA, uint32_t *__restrict__ sec,
                                                   uint32_t *__restrict__ sufl,
                                                   uint8_t *__restrict__ L, uint8_t *__restrict__ lrow,
                                                   const uint8_t *__restrict__ T,
                                                   const uint16_t *__restrict__ blk,
                                                   const BlockInfo *__restrict__ blocks, uint32_t n, uint32_t round,
                                                   uint32_t *__restrict__ nh, uint32_t *__restrict__ ns,
                                                   uint32_t *__restrict__ ncnt, uint32_t *__restrict__ tmax,
                                                   uint32_t rank_max) {
    // a tile's rows: the groups starting in its SEG_TILE rows, each of at most SEG_GROUP rows. CAP keeps the LDS of
    // a radix workgroup within a quarter of a CU's, so four share a CU (the rounds are latency-bound)
    constexpr int CAP = SEG_TILE + SEG_GROUP;
    constexpr int MAXG = SEG_TILE / 2;         // groups starting in the tile (two rows or more each)
    constexpr int ITEMS = CAP / 256;
    constexpr int W = 4;
    constexpr bool RADIX = MODE == 1;
    __shared__ uint32_t s_gh[MAXG], s_go[MAXG];
    __shared__ uint32_t s_hoff[MODE ? 1 : MAXG], s_gend[MODE ? 1 : MAXG];   // gather: each group's offset and block end
    __shared__ uint32_t s_first[RADIX ? 1 : MAXG];   // key of each group's first row; the ranking: whether it splits
    __shared__ uint32_t s_g0, s_g1, s_cnt, s_maxsz, s_skip;
    __shared__ uint64_t s_buf[RADIX ? 2 * CAP : 1];
    __shared__ uint32_t s_sec[MODE == 2 ? CAP : 1];
    __shared__ uint32_t s_i[MODE ? CAP : 1];
    __shared__ uint8_t s_lb[RADIX && CARRY ? CAP : 1];   // radix split: each row's BWT byte by its slot before the sort
    __shared__ uint32_t s_wcnt[RADIX ? W : 1][256];
    __shared__ uint32_t s_start[RADIX ? 256 : 1];
    __shared__ uint32_t s_scratch[16];
    __shared__ uint64_t s_or[W], s_and[W];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    if (tid == 0) {
        s_g0 = tstart[blockIdx.x];
        s_g1 = tstart[blockIdx.x + 1];
        s_maxsz = 0;
        if (MODE) {
            const uint32_t mx = tmax[blockIdx.x];
            s_skip = RADIX ? mx <= rank_max : mx > rank_max;   // the other split kernel takes this tile
        }
    }
    __syncthreads();
    if (MODE && s_skip) return;
    const uint32_t g0 = s_g0, g1 = s_g1, ng = g1 - g0;
    if (ng == 0) {
        if (tid == 0) {
            if (MODE == 0) tmax[blockIdx.x] = 0;
            else ncnt[blockIdx.x] = 0;
        }
        return;
    }
    const uint32_t span0 = go[g0], span1 = g1 < G ? go[g1] : E;
    const uint32_t cnt = span1 - span0;        // <= CAP - 1
    for (uint32_t j = tid; j < ng; j += 256) {
        const uint32_t h = gh[g0 + j];
        s_gh[j] = h;
        s_go[j] = go[g0 + j] - span0;
        if (MODE == 0) {   // a group lies in one block
            const BlockInfo bi = blocks[blk ? blk[h] : 0u];
            s_hoff[j] = round_offset(bi, round);
            s_gend[j] = bi.off + bi.n;
        }
    }
    __syncthreads();

    // gather: row e of the tile belongs to group lg (binary search over the tile's group offsets). MODE 0 reads each
    // row's suffix from SA, its BWT byte from L and its key at i + h, and stores them in list order (row span0 + e of
    // the round: sufl, lrow, sec), so the split reads them back contiguously. The ITEMS rows of a thread are independent, so their loads
    // are issued together.
    uint32_t ga[ITEMS], gi[ITEMS], gv[ITEMS];
    uint8_t gb[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t e = tid + k * 256;
        uint32_t a = 0, b = ng;                // last lg with s_go[lg] <= e
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (s_go[mid] <= e) a = mid; else b = mid;
        }
        ga[k] = a;
    }
    if (MODE == 0) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            const uint32_t r = s_gh[ga[k]] + (e - s_go[ga[k]]);   // an SA row of the tile's groups below cnt
            const uint32_t a = SA[e < cnt ? r : s_gh[0]];
            gb[k] = CARRY ? L[e < cnt ? r : s_gh[0]] : (uint8_t)0;
            gi[k] = e < cnt ? a : 0u;
        }
        // the key of suffix i at offset h inside its block [.., end): ISA[i + h] + h, or end - 1 - i past the end.
        // Every key's ISA load is in flight together: the index is clamped into the block (and row 0 past the tile's
        // rows) so that no load depends on a branch, and the past-the-end case is chosen afterwards
        uint32_t raw[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            const uint64_t j = (uint64_t)gi[k] + s_hoff[ga[k]];
            raw[k] = ISA[e < cnt ? (j < s_gend[ga[k]] ? (uint32_t)j : s_gend[ga[k]] - 1u) : 0u];
        }
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            const uint32_t h = s_hoff[ga[k]], end = s_gend[ga[k]];
            gv[k] = (uint64_t)gi[k] + h < end ? raw[k] + h : end - 1u - gi[k];
            if (e < cnt && e == s_go[ga[k]]) s_first[ga[k]] = gv[k];
        }
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            if (e < cnt) {
                sufl[span0 + e] = gi[k];
                if (CARRY) lrow[span0 + e] = gb[k];
                sec[span0 + e] = gv[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256, lg = ga[k];
            if (e < cnt && gv[k] != s_first[lg]) atomicMax(&s_maxsz, (lg + 1 < ng ? s_go[lg + 1] : cnt) - s_go[lg]);
        }
        __syncthreads();
        if (tid == 0) tmax[blockIdx.x] = s_maxsz;
        return;
    }
    // the rows' suffixes and keys, all loads in flight together: past the tile's rows the loads repeat its last row
    // (a load under a branch waits for everything before it)
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t e = tid + k * 256, ec = e < cnt ? e : cnt - 1;
        const uint32_t a = sufl[span0 + ec], v = sec[span0 + ec];
        gb[k] = CARRY ? lrow[span0 + ec] : (uint8_t)0;
        gi[k] = e < cnt ? a : 0u;
        gv[k] = e < cnt ? v : 0u;
    }

    // Direct ranking: each row counts, within its group, the rows of smaller sec (less) and of equal sec before it
    // (eqb). Its new SA position is head_of_group + less + eqb, its subgroup's head head_of_group + less, and the
    // last row of a subgroup of two or more emits it. Same order as the radix split (rows of equal sec keep their
    // slot order). Both write nothing for a group that does not split; a row that moves writes its suffix at its
    // new SA position (and in the first round its byte).
    if (MODE == 2) {
        // a group splits when some row's key differs from its first row's; the rows of a group that does not
        // split know their ranks without comparing
        for (uint32_t j = tid; j < ng; j += 256) s_first[j] = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            if (e < cnt) {
                s_sec[e] = gv[k];
                s_i[e] = gi[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            if (e < cnt && gv[k] != s_sec[s_go[ga[k]]]) s_first[ga[k]] = 1;
        }
        __syncthreads();
        uint32_t eh[ITEMS], es[ITEMS];
        uint32_t nemit = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = tid + k * 256;
            es[k] = 0;
            eh[k] = 0;
            if (e < cnt) {
                const uint32_t lg = ga[k], se = gv[k];
                const uint32_t st = s_go[lg], en = lg + 1 < ng ? s_go[lg + 1] : cnt;
                uint32_t less = 0, eq = en - st, eqb = e - st;
                if (s_first[lg]) {
                    eq = 0;
                    eqb = 0;
                    for (uint32_t j = st; j < en; j++) {
                        const uint32_t sj = s_sec[j];
                        less += sj < se ? 1u : 0u;
                        eq += sj == se ? 1u : 0u;
                        eqb += (sj == se && j < e) ? 1u : 0u;
                    }
                }
                const uint32_t i = gi[k];
                const uint32_t head = s_gh[lg] + less, pnew = head + eqb;
                if (eq != en - st) {             // the group splits (else SA, ISA stay as they are)
                    // only what changes: a row that keeps its SA position or its head (the first subgroup keeps
                    // the group's) writes nothing there
                    if (pnew != s_gh[lg] + (e - st)) {
                        SA[pnew] = i;
                        if (CARRY) L[pnew] = gb[k];
                    }
                    if (less) ISA[i] = head;
                    if (!CARRY && eq == 1) L[pnew] = i ? T[i - 1] : (uint8_t)0;
                }
                if (eq >= 2 && eqb == eq - 1) { eh[k] = head; es[k] = eq; nemit++; }
            }
        }
        uint32_t total_emit;
        const uint32_t my_emit = block_excl_sum<256>(nemit, s_scratch, &total_emit);
        if (tid == 0) ncnt[blockIdx.x] = total_emit;
        size_t at = (size_t)blockIdx.x * SEG_TILE + my_emit;
#pragma unroll
        for (int k = 0; k < ITEMS; k++)
            if (es[k]) { nh[at] = eh[k]; ns[at] = es[k]; at++; }
        return;
    }

    // The sort key packs (group, value, row): the row's slot in the low 12 bits (carried, not sorted), above it a
    // value that orders and ties the rows of a group as their secs do, and the group right above it. The value is the
    // sec itself with only the bits that vary over the tile (the bits above the highest varying one are equal in every
    // row, so dropping them keeps the order), or the sec minus its group's smallest one when that needs fewer bits:
    // on repetitive text the secs of one group lie close together while the tile's span the whole rank range. Only the
    // digits of the value and group bits are sorted.
    uint32_t *const s_gmin = &s_wcnt[0][0];   // the radix counters are free until the first pass clears them
    static_assert(W * 256 >= MAXG, "group minima in the counters' LDS");
    for (uint32_t j = tid; j < ng; j += 256) s_gmin[j] = ~0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        if (tid + k * 256 < cnt) atomicMin(&s_gmin[ga[k]], gv[k]);
    __syncthreads();
    uint32_t so = 0, sa = ~0u, rmax = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t e = tid + k * 256;
        if (e < cnt) {
            so |= gv[k];
            sa &= gv[k];
            rmax = max(rmax, gv[k] - s_gmin[ga[k]]);
        }
    }
    for (int o = 1; o < WAVE; o <<= 1) {
        so |= __shfl_xor(so, o, WAVE);
        sa &= __shfl_xor(sa, o, WAVE);
        rmax = max(rmax, (uint32_t)__shfl_xor(rmax, o, WAVE));
    }
    __shared__ uint32_t s_rmax[W];
    if (lane == 0) { s_or[wave] = so; s_and[wave] = sa; s_rmax[wave] = rmax; }
    __syncthreads();
    const uint32_t svary =
        (uint32_t)((s_or[0] | s_or[1] | s_or[2] | s_or[3]) ^ (s_and[0] & s_and[1] & s_and[2] & s_and[3]));
    const uint32_t rtop = max(max(s_rmax[0], s_rmax[1]), max(s_rmax[2], s_rmax[3]));
    const uint32_t tbits = svary ? 32u - (uint32_t)__clz(svary) : 0u;               // varying sec bits of the tile
    const uint32_t rbits = rtop ? 32u - (uint32_t)__clz(rtop) : 0u;                 // bits of the widest group span
    const bool rel = rbits < tbits;
    const uint32_t sbits = rel ? rbits : tbits;                                      // value bits
    const uint32_t gbits = ng > 1 ? 32u - (uint32_t)__clz(ng - 1u) : 0u;            // group bits
    const uint64_t smask = (1ull << sbits) - 1ull;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t e = tid + k * 256;
        if (e < cnt) {
            const uint32_t val = rel ? gv[k] - s_gmin[ga[k]] : gv[k];
            s_buf[e] = (((uint64_t)ga[k] << sbits) | ((uint64_t)val & smask)) << 12 | e;
            s_i[e] = gi[k];
            if (CARRY) s_lb[e] = gb[k];
        }
    }
    __syncthreads();
    const uint32_t kbits = sbits + gbits;   // sorted bits above the row index

    // LDS radix sort of the packed keys over the 8-bit digits of the kept bits (stable multisplit per pass). Each
    // wave ranks a contiguous run of ipt * 64 slots, ipt = ceil(cnt / 256): the rows spread evenly over the waves
    int src = 0;
    const uint32_t ipt = (cnt + 255) / 256;
    for (uint32_t dsh = 0; dsh < kbits; dsh += 8) {
        const uint32_t shift = 12 + dsh;
        for (int i = tid; i < W * 256; i += 256) (&s_wcnt[0][0])[i] = 0;
        __syncthreads();
        uint64_t v[ITEMS];
        uint32_t rk[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            v[k] = 0;
            rk[k] = 0;
            if ((uint32_t)k < ipt) {   // uniform over the wave
                const uint32_t slot = (wave * ipt + k) * WAVE + lane;
                const bool valid = slot < cnt;
                v[k] = valid ? s_buf[src * CAP + slot] : 0ull;
                const uint32_t d = (uint32_t)(v[k] >> shift) & 255u;
                const uint64_t peers = match_digit8(d, __ballot(valid));
                if (valid) {
                    const uint32_t before = count_below(peers);
                    const uint32_t c = s_wcnt[wave][d];
                    rk[k] = c + before;
                    s_wcnt[wave][d] = c + (uint32_t)__popcll(peers);
                }
            }
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w2 = 0; w2 < W; w2++) { uint32_t c = s_wcnt[w2][tid]; s_wcnt[w2][tid] = tot; tot += c; }
        const uint32_t st = block_excl_sum<256>(tot, s_scratch, nullptr);
        s_start[tid] = st;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t slot = (wave * ipt + k) * WAVE + lane;
            if ((uint32_t)k < ipt && slot < cnt) {
                const uint32_t d = (uint32_t)(v[k] >> shift) & 255u;
                s_buf[(src ^ 1) * CAP + s_start[d] + s_wcnt[wave][d] + rk[k]] = v[k];
            }
        }
        __syncthreads();
        src ^= 1;
    }

    // split the groups: blocked rows per thread; a subgroup starts where (group, sec) changes, its head is the SA
    // position of its first slot (max-scan over slots: the groups' heads themselves come in any order), and each
    // kept subgroup's last slot emits it into the next list
    constexpr int PER = CAP / 256;
    uint32_t pk[PER], hk[PER], ik[PER], gk[PER], ek[PER];
    bool flag[PER], lastk[PER];
    uint32_t run_max = 0, nemit = 0;
    const uint64_t *buf = s_buf + src * CAP;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t slot = tid * PER + k;
        flag[k] = false;
        lastk[k] = false;
        pk[k] = 0;
        ik[k] = 0;
        gk[k] = 0;
        ek[k] = 0;
        if (slot < cnt) {
            const uint64_t key = buf[slot] >> 12;
            const uint32_t lg = (uint32_t)(key >> sbits);
            gk[k] = lg;
            pk[k] = s_gh[lg] + (slot - s_go[lg]);
            ek[k] = (uint32_t)(buf[slot] & 0xFFFu);   // the row's slot before the sort
            ik[k] = s_i[ek[k]];
            flag[k] = slot == 0 || (buf[slot - 1] >> 12) != key;
            lastk[k] = slot + 1 == cnt || (buf[slot + 1] >> 12) != key;
        }
        const uint32_t cand = flag[k] ? slot : 0u;
        run_max = run_max > cand ? run_max : cand;
        hk[k] = run_max;   // within-thread running max of subgroup-start slots, fixed below with earlier threads
    }
    uint32_t *const s_nsub = &s_wcnt[0][0];   // the radix counters are free after the last pass
    static_assert(W * 256 >= MAXG, "subgroup counts in the counters' LDS");
    for (uint32_t j = tid; j < ng; j += 256) s_nsub[j] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++)
        if (flag[k]) atomicAdd(&s_nsub[gk[k]], 1u);
    uint32_t ex_s, in_m, tot_s, tot_m;
    // the max over earlier threads: inclusive max of each thread's run_max, shifted by one thread
    block_scan_sum_max<256>(0u, run_max, ex_s, in_m, tot_s, tot_m, s_scratch);
    const uint32_t prev_max = __shfl_up(in_m, 1, WAVE);
    __shared__ uint32_t s_wmax[W];
    if (lane == WAVE - 1) s_wmax[wave] = in_m;
    __syncthreads();
    uint32_t before_me = lane > 0 ? prev_max : (wave > 0 ? s_wmax[wave - 1] : 0u);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t slot = tid * PER + k;
        if (slot < cnt) {
            const uint32_t hslot = hk[k] > before_me ? hk[k] : before_me;
            const uint32_t head = s_gh[gk[k]] + (hslot - s_go[gk[k]]);
            hk[k] = head;
            if (s_nsub[gk[k]] > 1) {   // only what changes (see the direct ranking above)
                if (pk[k] != s_gh[gk[k]] + (ek[k] - s_go[gk[k]])) {
                    SA[pk[k]] = ik[k];
                    if (CARRY) L[pk[k]] = s_lb[ek[k]];
                }
                if (head != s_gh[gk[k]]) ISA[ik[k]] = head;
                if (!CARRY && flag[k] && lastk[k]) L[pk[k]] = ik[k] ? T[ik[k] - 1] : (uint8_t)0;
            }
            if (lastk[k] && !flag[k]) nemit++;
        }
    }
    uint32_t total_emit;
    const uint32_t my_emit = block_excl_sum<256>(nemit, s_scratch, &total_emit);
    if (tid == 0) ncnt[blockIdx.x] = total_emit;
    size_t at = (size_t)blockIdx.x * SEG_TILE + my_emit;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t slot = tid * PER + k;
        if (slot < cnt && lastk[k] && !flag[k]) {
            nh[at] = hk[k];
            ns[at] = pk[k] - hk[k] + 1;
            at++;
        }
    }
}

// The emissions of k_seg_round's tiles (tile t: ecnt[t] entries at t * SEG_TILE) packed after the prefix eoff[t].
__global__ __launch_bounds__(256) void k_compact_emits(const uint32_t *__restrict__ eh, const uint32_t *__restrict__ es,
                                                       const uint32_t *__restrict__ ecnt,
                                                       const uint32_t *__restrict__ eoff, uint32_t *__restrict__ nh,
                                                       uint32_t *__restrict__ ns) {
    const uint32_t t = blockIdx.x, c = ecnt[t], o = eoff[t];
    const size_t src = (size_t)t * SEG_TILE;
    for (uint32_t k = threadIdx.x; k < c; k += 256) {
        nh[o + k] = eh[src + k];
        ns[o + k] = es[src + k];
    }
}

// ------------------------------------------------------------------ launch wrappers

void launch_scan_groups(const uint32_t *gs, uint32_t G, uint32_t *go, uint32_t *tot, uint32_t *tstart,
                        uint64_t *status, uint32_t *counter, uint32_t epoch, hipStream_t s) {
    const uint32_t ntiles = (G + 2047) / 2048;
    if (ntiles) k_scan_groups<<<ntiles, 256, 0, s>>>(gs, G, go, tot, tstart, status, counter, epoch, ntiles);
}
void launch_compact_emits(const uint32_t *eh, const uint32_t *es, const uint32_t *ecnt, const uint32_t *eoff,
                          uint32_t ntiles, uint32_t *nh, uint32_t *ns, hipStream_t s) {
    if (ntiles) k_compact_emits<<<ntiles, 256, 0, s>>>(eh, es, ecnt, eoff, nh, ns);
}
void launch_large_elems(const uint32_t *lh, const uint32_t *lo, uint32_t G, uint32_t m, uint32_t *pos, uint32_t *grp,
                        hipStream_t s) {
    uint32_t grid = (m + 255) / 256;
    if (grid > 4096) grid = 4096;
    if (grid) k_large_elems<<<grid, 256, 0, s>>>(lh, lo, G, m, pos, grp);
}
void launch_seg_round(int mode, const SegRoundArgs &a, hipStream_t s) {
    const uint32_t ntiles = (a.E + SEG_TILE - 1) / SEG_TILE;
    if (!ntiles) return;
#define KBWT_SEG_ARGS a.gh, a.go, a.tstart, a.G, a.E, a.SA, a.ISA, a.sec, a.sufl, a.L, a.lrow, a.T, a.blk, a.blocks, \
                      a.n, a.round, a.nh, a.ns, a.ncnt, a.tmax, a.rank_max
    // every tile goes to exactly one of the two split kernels (modes 1 and 2)
    const bool carry = a.round == 1;
    if (mode == 0) {
        if (carry) k_seg_round<0, true><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
        else k_seg_round<0, false><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
    } else if (mode == 1) {
        if (carry) k_seg_round<1, true><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
        else k_seg_round<1, false><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
    } else {
        if (carry) k_seg_round<2, true><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
        else k_seg_round<2, false><<<ntiles, 256, 0, s>>>(KBWT_SEG_ARGS);
    }
#undef KBWT_SEG_ARGS
}

}  // namespace kbwt
