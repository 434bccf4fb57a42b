// Device-side building blocks shared by the forward and inverse kernels (gfx950, wave64).
//
//  - decoupled-lookback status words: one 64-bit word per (tile, lane of interest) holding an epoch, a flag
//    (aggregate or inclusive prefix) and a 32-bit value. The epoch makes a status array reusable by every launch
//    without clearing it: a word written by an earlier launch carries an older epoch and reads as "not ready".
//  - stable wave-level multisplit ranking of 8-bit digits with ballots (wave64 peers), the core of the radix passes
//    and of the LF-map construction.
//  - block scans for 256- or 512-thread workgroups.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kbwt {

constexpr int WAVE = 64;
constexpr uint32_t ST_AGG = 1u, ST_INC = 2u;
constexpr uint32_t EPOCH_MASK = 0x3FFFFFFFu;

__device__ __forceinline__ uint64_t st_pack(uint32_t epoch, uint32_t flag, uint32_t v) {
    return ((uint64_t)(epoch & EPOCH_MASK) << 34) | ((uint64_t)flag << 32) | (uint64_t)v;
}
__device__ __forceinline__ void st_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t st_load(const uint64_t *p) {
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix over tiles [0, tile) of the quantity whose status words sit at status[t * stride]. Sum or max.
// Called by one thread per quantity. Every earlier tile has taken its tile id before this one (ids come from an
// atomic counter), publishes its aggregate before it looks back itself, and so every wait here ends.
template <bool MAX>
__device__ __forceinline__ uint32_t lookback(const uint64_t *status, size_t stride, uint32_t tile, uint32_t epoch) {
    uint32_t acc = 0;
    int64_t t = (int64_t)tile - 1;
    const uint32_t want = epoch & EPOCH_MASK;
    while (t >= 0) {
        uint64_t s = st_load(status + (size_t)t * stride);
        uint32_t e = (uint32_t)(s >> 34), f = (uint32_t)(s >> 32) & 3u;
        if (e != want || f == 0) { __builtin_amdgcn_s_sleep(1); continue; }
        uint32_t v = (uint32_t)s;
        acc = MAX ? (acc > v ? acc : v) : acc + v;
        if (f == ST_INC) break;
        --t;
    }
    return acc;
}

// The same with a window: one thread reads the status words of LB earlier tiles per round trip, adds the ready
// aggregates from the nearest one on, stops at the first inclusive prefix and re-reads from the first word that was
// not ready yet.
template <int LB>
__device__ __forceinline__ uint32_t lookback_win(const uint64_t *status, size_t stride, uint32_t tile,
                                                 uint32_t epoch) {
    uint32_t acc = 0;
    int64_t t = (int64_t)tile - 1;
    const uint32_t want = epoch & EPOCH_MASK;
    while (t >= 0) {
        uint64_t w[LB];
#pragma unroll
        for (int j = 0; j < LB; j++) w[j] = (t - j >= 0) ? st_load(status + (size_t)(t - j) * stride) : 0ull;
        uint32_t add = 0;
        int adv = 0;
        bool stop = false, done = false;
#pragma unroll
        for (int j = 0; j < LB; j++) {
            if (!stop) {
                if (t - j < 0) {
                    stop = true;
                    done = true;
                } else {
                    uint32_t e = (uint32_t)(w[j] >> 34), f = (uint32_t)(w[j] >> 32) & 3u;
                    if (e != want || f == 0) stop = true;
                    else {
                        add += (uint32_t)w[j];
                        adv = j + 1;
                        if (f == ST_INC) { stop = true; done = true; }
                    }
                }
            }
        }
        acc += add;
        t -= adv;
        if (done) break;
        if (adv == 0) __builtin_amdgcn_s_sleep(1);
    }
    return acc;
}

// Wave-wide lookback of one quantity (sum or max): lane j reads the word of tile t-1-j, so one round trip covers 64
// earlier tiles. Every lane of the wave must call it; all get the result.
template <bool MAX>
__device__ __forceinline__ uint32_t lookback_wave(const uint64_t *status, uint32_t tile, uint32_t epoch) {
    const uint32_t lane = __lane_id();
    const uint32_t want = epoch & EPOCH_MASK;
    uint32_t acc = 0;
    int64_t t = (int64_t)tile - 1;
    while (t >= 0) {
        int64_t mine = t - (int64_t)lane;
        uint64_t w = mine >= 0 ? st_load(status + mine) : 0ull;
        uint32_t e = (uint32_t)(w >> 34), f = (uint32_t)(w >> 32) & 3u;
        bool ready = mine >= 0 && e == want && f != 0;
        bool inc = ready && f == ST_INC;
        uint64_t notready = ~__ballot(ready);
        uint64_t incs = __ballot(inc);
        // lanes that count: those before the first lane that is not ready, and up to the first inclusive one
        uint32_t first_nr = notready ? (uint32_t)__builtin_ctzll(notready) : 64u;
        uint32_t first_inc = incs ? (uint32_t)__builtin_ctzll(incs) : 64u;
        bool done = first_inc < first_nr;
        uint32_t upto = done ? first_inc + 1 : first_nr;
        uint32_t v = lane < upto ? (uint32_t)w : 0u;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            uint32_t u = __shfl_xor(v, o, WAVE);
            v = MAX ? (v > u ? v : u) : v + u;
        }
        acc = MAX ? (acc > v ? acc : v) : acc + v;
        if (done) break;
        if (upto == 0) __builtin_amdgcn_s_sleep(1);
        t -= upto;
    }
    return acc;
}

// Tile id from a per-launch counter; the workgroup that draws the last id resets it for the next launch (all
// ntiles draws have happened by then).
__device__ __forceinline__ uint32_t next_tile(uint32_t *counter, uint32_t ntiles) {
    uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == ntiles - 1) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
}

// Lanes of mask below this lane (mbcnt: two instructions, no lane mask to build).
__device__ __forceinline__ uint32_t count_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Bit B of d spread over a word (all ones when set): one signed bit-field extract. The compiler turns the equivalent
// shift pair back into two shifts and a sign test, so the instruction is written out.
template <int B> __device__ __forceinline__ uint32_t bit_mask(uint32_t d) {
    uint32_t s;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(s) : "v"(d), "i"(B));
    return s;
}
// acc | (m ^ s) in one three-input logic operation: truth table 0xF6 over (acc, m, s), indexed acc*4 + m*2 + s.
__device__ __forceinline__ uint32_t or_xor(uint32_t acc, uint32_t m, uint32_t s) {
    return __builtin_amdgcn_bitop3_b32(acc, m, s, (unsigned char)0xF6);
}
template <int B> __device__ __forceinline__ void match_bit(uint32_t d, uint32_t &dlo, uint32_t &dhi) {
    const uint32_t s = bit_mask<B>(d);
    const uint64_t m = __ballot(s != 0u);
    dlo = or_xor(dlo, (uint32_t)m, s);
    dhi = or_xor(dhi, (uint32_t)(m >> 32), s);
}
// Peers of this lane: the valid lanes whose 8-bit digit equals d. valid_mask = ballot(valid). One ballot per digit
// bit; the lanes that differ from this one in bit b are that ballot XOR the lane's own bit spread over a word, and
// the peers are the valid lanes that differ in no bit (an extract, a compare and two three-input logic operations per
// bit).
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, uint64_t valid_mask) {
    uint32_t dlo = 0, dhi = 0;
    match_bit<0>(d, dlo, dhi);
    match_bit<1>(d, dlo, dhi);
    match_bit<2>(d, dlo, dhi);
    match_bit<3>(d, dlo, dhi);
    match_bit<4>(d, dlo, dhi);
    match_bit<5>(d, dlo, dhi);
    match_bit<6>(d, dlo, dhi);
    match_bit<7>(d, dlo, dhi);
    return valid_mask & ~(((uint64_t)dhi << 32) | dlo);
}

// Wave inclusive scans with shuffles.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t u = __shfl_up(v, o, WAVE);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t u = __shfl_up(v, o, WAVE);
        if (lane >= (uint32_t)o) v = v > u ? v : u;
    }
    return v;
}

// Block-wide exclusive sum and inclusive max of one value per thread, in thread order. scratch holds 2 * waves
// words. Returns the block totals too. Contains barriers: every thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ void block_scan_sum_max(uint32_t vs, uint32_t vm, uint32_t &ex_sum, uint32_t &in_max,
                                                   uint32_t &tot_sum, uint32_t &tot_max, uint32_t *scratch) {
    constexpr int W = BLOCK / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    uint32_t is = wave_incl_sum(vs), im = wave_incl_max(vm);
    if (lane == WAVE - 1) { scratch[wave] = is; scratch[W + wave] = im; }
    __syncthreads();
    uint32_t ps = 0, pm = 0, ts = 0, tm = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        uint32_t s = scratch[w], m = scratch[W + w];
        if (w < wave) { ps += s; pm = pm > m ? pm : m; }
        ts += s; tm = tm > m ? tm : m;
    }
    __syncthreads();
    ex_sum = ps + is - vs;
    in_max = pm > im ? pm : im;
    tot_sum = ts; tot_max = tm;
}

// Block-wide exclusive sum (one value per thread). REUSE: end with a barrier so that scratch can be written again
// right away; a caller that never writes it again in the launch passes false.
template <int BLOCK, bool REUSE = true>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t *scratch, uint32_t *total) {
    constexpr int W = BLOCK / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    uint32_t is = wave_incl_sum(v);
    if (lane == WAVE - 1) scratch[wave] = is;
    __syncthreads();
    uint32_t p = 0, t = 0;
#pragma unroll
    for (int w = 0; w < W; w++) { uint32_t s = scratch[w]; if (w < wave) p += s; t += s; }
    if (REUSE) __syncthreads();
    if (total) *total = t;
    return p + is - v;
}

}  // namespace kbwt
