// The library's CPU path (kbwt_cpu.cpp): the forward and inverse BWT on host threads, with the single-block
// functions' conventions and results. Used only as the opt-in fallback after a device error. Returns 0, or -2 when
// host memory runs out. Arguments are validated by the callers; T == U is allowed.
#pragma once
#include <stdint.h>

namespace kbwt {

int32_t cpu_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I, int32_t threads);
int32_t cpu_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, const int32_t *I, int32_t threads);

}  // namespace kbwt
