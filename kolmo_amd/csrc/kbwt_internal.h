// Internal interface between the kernels (kbwt_kernels.hip) and the host orchestration (kbwt_host.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace kbwt {

constexpr int MAXP = 8;                      // 8-bit digits of a 64-bit key
constexpr int SORT_BLOCK = 256, SORT_ITEMS = 16, SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int RG_BLOCK = 256, RG_ITEMS = 8, RG_TILE = RG_BLOCK * RG_ITEMS;
constexpr int LF_BLOCK = 256, LF_ITEMS = 32, LF_TILE = LF_BLOCK * LF_ITEMS;
constexpr int HIST_GRID = 512;               // workgroups of the histogram kernels

struct Shifts {
    uint32_t P;
    uint32_t s[MAXP];
};

struct SortArgs {
    const uint64_t *kin;
    const uint32_t *vin;
    uint64_t *kout;
    uint32_t *vout;
    uint32_t m, shift;
    const uint32_t *dbase;
    uint64_t *status;
    uint32_t *counter;
    uint32_t epoch;
    bool from_t;
    const uint8_t *T, *code;
    uint32_t w, h0;
};

struct RegroupArgs {
    const uint64_t *K;
    const uint32_t *V, *pos;
    uint32_t m, kshift;
    uint32_t *SA, *ISA;
    uint8_t *L;
    const uint8_t *T;
    uint32_t *pos_out, *grp_out, *m_out;
    uint64_t *st_sum, *st_max;
    uint32_t *counter, epoch;
};

void launch_byte_hist(const uint8_t *T, uint32_t n, uint32_t *counts, uint32_t grid, hipStream_t s);
void launch_scan_digits(uint32_t *counts, uint32_t nsets, uint32_t *tot, uint32_t *excl, uint32_t base,
                        hipStream_t s);
void launch_hist_T(const uint8_t *T, uint32_t n, const uint8_t *code, uint32_t w, uint32_t h0, uint32_t P,
                   uint32_t *counts, uint32_t grid, hipStream_t s);
void launch_hist_keys(const uint64_t *K, uint32_t m, const Shifts &sh, uint32_t *counts, uint32_t grid,
                      hipStream_t s);
void launch_sort_pass(const SortArgs &a, hipStream_t s);
void launch_regroup(const RegroupArgs &a, hipStream_t s);
void launch_round_keys(const uint32_t *pos, const uint32_t *grp, uint32_t m, const uint32_t *SA, const uint32_t *ISA,
                       uint32_t n, uint32_t h, uint64_t *K, uint32_t *V, unsigned long long *orand, hipStream_t s);
void launch_bwt(const uint8_t *T, const uint8_t *L, const uint32_t *ISA, uint8_t *U, uint32_t n, hipStream_t s);
void launch_samples(const uint32_t *ISA, uint32_t r, uint32_t count, int32_t *I, hipStream_t s);
void launch_lf_pass(const uint8_t *L, uint32_t n, const uint32_t *cbase, uint32_t *lf, uint64_t *status,
                    uint32_t *counter, uint32_t epoch, hipStream_t s);
void launch_walk_count(const uint32_t *lf, uint32_t n, uint32_t P, uint32_t dshift, uint32_t nreg, uint32_t S,
                       const uint32_t *cbase, uint64_t *link, uint32_t *len, uint8_t *stage, uint32_t cap,
                       uint32_t *cont, hipStream_t s);
void launch_rank_jump(const uint64_t *in, uint64_t *out, uint32_t S, hipStream_t s);
void launch_walk_emit(const uint32_t *lf, uint32_t n, uint32_t P, uint32_t S, const uint64_t *link,
                      const uint32_t *len, const uint32_t *cbase, const uint8_t *stage, uint32_t cap,
                      const uint32_t *cont, uint8_t *out, hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s);

}  // namespace kbwt
