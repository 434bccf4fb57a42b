// Internal interface between the kernels (kbwt_kernels.hip) and the host orchestration (kbwt_host.hip).
//
// Every call works on a batch of B blocks packed back to back: block b owns bytes, suffixes and suffix-array rows
// [off, off + n) of the packed index space. A single-block call is the batch B = 1 (blk == nullptr).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace kbwt {

constexpr int MAXP = 8;                      // 8-bit digits of a 64-bit key
constexpr int RG_BLOCK = 256, RG_ITEMS = 8, RG_TILE = RG_BLOCK * RG_ITEMS;
constexpr int LF_BLOCK = 256, LF_ITEMS = 32, LF_TILE = LF_BLOCK * LF_ITEMS;
constexpr int HIST_TILE = 4096;              // bytes per tile of the histogram kernels
constexpr uint32_t GRAM_UNIT = 1u << 17;     // bytes per workgroup of the m-gram histograms
constexpr uint32_t GRAM_MAXBITS = 14;        // widest m-gram (bins of one LDS table)

// An 8-bit digit of a first key, counted from the key's bottom, overlaps at most gram_len(w) consecutive codes of w
// bits. Blocks with gram_len(w) * w <= GRAM_MAXBITS (every width but 5 bits) take the m-gram digit histograms.
__host__ __device__ inline uint32_t gram_len(uint32_t w) {
    constexpr uint8_t len[9] = {8, 8, 4, 4, 2, 3, 2, 2, 1};
    return w <= 8 ? len[w] : 8;
}
__host__ __device__ inline bool gram_block(uint32_t w) { return gram_len(w) * w <= GRAM_MAXBITS; }
constexpr int HIST_GRID = 512;               // workgroups of k_hist_keys
constexpr int WALK_TILE = 256;               // splitters per workgroup of the inverse walks
constexpr uint32_t MIN_TILE = 2048;          // smallest tile of any tiled kernel (sizes the lookback arrays)
constexpr int SEG_TILE = 1024;               // doubling rounds: rows of the group list per tile (by first row)
constexpr int SEG_GROUP = 512;               // groups of up to SEG_GROUP rows sort inside one workgroup
constexpr int SEG_RANK_MAX = 64;             // default: tiles whose groups all have at most this many rows rank directly
                                             // (O(size) comparisons per row beat two to four LDS radix passes up
                                             // to 64 rows; at 256 the large groups' quadratic cost dominates)

// One block of a batch as the kernels see it.
struct BlockInfo {
    uint32_t off;      // first packed index of the block
    uint32_t n;        // length
    uint32_t w, h0;    // forward: bits of the symbol code, symbols in the first key
    uint32_t r;        // forward: sampling interval
    uint32_t soff;     // forward: first entry of the block's sampled indexes
    uint32_t P;        // inverse: primary index (the end marker's row, 1..n)
    uint32_t roff;     // inverse: first global row (off + block number; a block has n + 1 rows)
    uint32_t spoff;    // inverse: first splitter
    uint32_t nreg;     // inverse: splitters at rows 0, D, 2D, ... of the block
};

// Tiles of a tiled kernel over the blocks of a batch: block b owns tiles [first[b], first[b+1]). A kernel over one
// flat range of m elements uses first == nullptr.
struct Tiling {
    const uint32_t *first;
    const BlockInfo *blocks;
    uint32_t B;
};

// Offset of doubling round r >= 1 in block bi: h0 * 2^(r-1) symbols, its sorted prefix after the round before (at
// most n: past the block's end every suffix is told apart by its length). Each block doubles from its own h0, so a
// block's rounds, and the share of its suffixes each round still sorts, do not depend on the rest of its batch.
__host__ __device__ inline uint32_t round_offset(const BlockInfo &bi, uint32_t r) {
    const uint64_t h = r - 1 < 40 ? (uint64_t)bi.h0 << (r - 1) : (uint64_t)bi.n;
    return h < bi.n ? (uint32_t)h : bi.n;
}

struct Shifts {
    uint32_t P;
    uint32_t s[MAXP];
};

struct SortCfg {
    int block, items, window;
};
constexpr int SORT_NCFG = 11;
extern const SortCfg SORT_CFGS[SORT_NCFG];

struct SortArgs {
    int cfg;
    const uint64_t *kin;
    const uint32_t *vin;
    uint64_t *kout;
    uint32_t *vout;
    uint32_t m, shift;
    const uint32_t *dbase;        // digit bases: dbase[b * dstride + d] (per block), or dbase[d]
    uint32_t dstride;
    Tiling tiling;                // per-block segmented sort when tiling.first != nullptr
    uint64_t *status;
    uint32_t *counter;
    uint32_t epoch;               // the lookback epoch, or its offset from *depoch when depoch is set
    const uint32_t *depoch;       // device epoch of a launch sequence a graph may replay (nullptr: host epochs)
    bool from_t;                  // first pass: keys from T with each block's code table
    const uint8_t *T, *codes;
};

struct RegroupArgs {
    const uint64_t *K;
    const uint32_t *V, *pos;
    uint32_t m, kshift;
    const uint16_t *blk;          // block of every packed index, or nullptr for one block
    const BlockInfo *blocks;      // with blk: the blocks (their first rows bound the carry search)
    uint32_t *SA, *ISA;
    uint64_t *pairs;              // when set: (suffix << 32 | head) per row instead of the ISA writes
    uint8_t *L;
    const uint8_t *T;
    uint32_t *carry;              // one word per RG_TILE rows: the head of the group holding the tile's first row
    uint32_t *sh, *ss, *lh, *ls;  // the next round's small and large group lists (head, size)
    uint32_t *cnt;                // their counters: cnt[0] small, cnt[1] large
};

struct SegRoundArgs {
    const uint32_t *gh, *go;      // the small groups: heads and first-row offsets (prefix sums of sizes)
    const uint32_t *tstart;       // first group of every SEG_TILE-row tile, and G at the end
    uint32_t G, E;                // groups, rows
    uint32_t *SA, *ISA, *sec;     // sec: each row's key, in list order (row go[g] + k of group g)
    uint32_t *sufl;               // each row's suffix, in list order
    uint8_t *L;
    uint8_t *lrow;                // first round: each row's BWT byte, in list order (carried from L)
    const uint8_t *T;
    const uint16_t *blk;
    const BlockInfo *blocks;      // every block (blocks[0] when blk == nullptr)
    uint32_t n, round;            // doubling round >= 1: offset round_offset(block, round)
    uint32_t *nh, *ns, *ncnt;     // the split's emissions: per tile, SEG_TILE entries at tile * SEG_TILE, and counts
    uint32_t *tmax;               // per tile: rows of its largest group (written by the gather)
    uint32_t rank_max;            // tiles whose groups have at most rank_max rows rank them directly
};

void launch_scan_groups(const uint32_t *gs, uint32_t G, uint32_t *go, uint32_t *tot, uint32_t *tstart,
                        uint64_t *status, uint32_t *counter, uint32_t epoch, hipStream_t s);
void launch_compact_emits(const uint32_t *eh, const uint32_t *es, const uint32_t *ecnt, const uint32_t *eoff,
                          uint32_t ntiles, uint32_t *nh, uint32_t *ns, hipStream_t s);
void launch_large_elems(const uint32_t *lh, const uint32_t *lo, uint32_t G, uint32_t m, uint32_t *pos, uint32_t *grp,
                        hipStream_t s);
void launch_seg_round(int mode, const SegRoundArgs &a, hipStream_t s);

void launch_byte_hist(const uint8_t *T, Tiling tl, uint32_t ntiles, uint32_t *counts, hipStream_t s);
void launch_scan_digits(uint32_t *counts, uint32_t nsets, uint32_t sets_per_block, const BlockInfo *blocks,
                        int base_mode, uint32_t *tot, uint32_t *excl, hipStream_t s);
void launch_hist_T(const uint8_t *T, Tiling tl, uint32_t ntiles, const uint8_t *codes, uint32_t P, uint32_t *counts,
                   hipStream_t s);
void launch_hist_gram(const uint8_t *T, Tiling units, uint32_t nunits, const uint8_t *codes, const uint32_t *tot,
                      uint32_t P, uint32_t *counts, hipStream_t s);
void launch_hist_keys(const uint64_t *K, uint32_t m, const Shifts &sh, uint32_t *counts, uint32_t grid,
                      hipStream_t s);
uint32_t sort_tile(int cfg);
void launch_sort_pass(const SortArgs &a, uint32_t ntiles, hipStream_t s);
void launch_regroup(const RegroupArgs &a, hipStream_t s);
constexpr int PART_BLOCK = 1024, PART_ITEMS = 16, PART_TILE = PART_BLOCK * PART_ITEMS;   // the ISA partition pass
void launch_partition_pairs(const uint64_t *in, uint64_t *out, uint32_t N, const Tiling &tl, uint32_t ntiles,
                            uint64_t *status, uint32_t *counter, uint32_t epoch, const uint32_t *depoch,
                            hipStream_t s);
void launch_isa_scatter(const uint64_t *pairs, uint32_t N, uint32_t maxn, const uint16_t *blk,
                        const BlockInfo *blocks, uint32_t *ISA, hipStream_t s);
void launch_round_keys(const uint32_t *pos, const uint32_t *grp, uint32_t m, const uint32_t *SA, const uint32_t *ISA,
                       uint32_t n, const uint16_t *blk, const BlockInfo *blocks, uint32_t round, uint64_t *K, uint32_t *V,
                       unsigned long long *orand, hipStream_t s);
void launch_bwt(const uint8_t *T, const uint8_t *L, const uint32_t *ISA, const uint16_t *blk, const BlockInfo *blocks,
                uint8_t *U, uint32_t N, hipStream_t s);
void launch_samples(const uint32_t *ISA, const BlockInfo *blocks, uint32_t B, uint32_t count, int32_t *I,
                    hipStream_t s);
void launch_fill_blk(uint16_t *blk, Tiling tl, uint32_t ntiles, hipStream_t s);   // tiles of HIST_TILE indexes
void launch_lf_pass(const uint8_t *L, Tiling tl, uint32_t ntiles, const uint32_t *cbase, uint32_t *lf,
                    uint64_t *status, uint32_t *counter, uint32_t epoch, const uint32_t *depoch, hipStream_t s);
// Lookback epochs: the host numbers launches in [1, EPOCH_HALF); a captured graph reads a device counter in
// [EPOCH_HALF, 2 * EPOCH_HALF) that a bump kernel at its start moves on by EPOCH_STRIDE.
constexpr uint32_t EPOCH_HALF = 1u << 29, EPOCH_STRIDE = 64;
void launch_epoch_bump(uint32_t *depoch, hipStream_t s);
void launch_lf_pairs(const uint32_t *lf, const uint8_t *L, Tiling tl, uint32_t ntiles, uint32_t dshift,
                     uint64_t *pair, hipStream_t s);
void launch_walk_count(const uint64_t *pair, Tiling tl, uint32_t ntiles, uint32_t dshift, uint32_t S,
                       const uint32_t *cbase, uint64_t *link, uint32_t *len, uint8_t *stage, size_t sstride,
                       uint32_t cap, uint32_t *cont, hipStream_t s);
uint32_t rank_jump_iters(uint32_t S);
void launch_rank_jump(const uint64_t *in, uint64_t *out, uint32_t S, hipStream_t s);
void launch_check_lists(const uint64_t *link, const BlockInfo *blocks, uint32_t B, uint32_t S, uint32_t *bad,
                        hipStream_t s);
void launch_walk_emit(const uint32_t *lf, Tiling tl, uint32_t ntiles, uint32_t dshift, uint32_t S,
                      const uint64_t *link, const uint32_t *len, const uint32_t *cbase, const uint8_t *stage,
                      size_t sstride, uint32_t cap, const uint32_t *cont, const uint32_t *bad, uint8_t *out,
                      hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s);

}  // namespace kbwt
