// Learned router: sizes and launchers shared by kbwt_router.hip (kernels), kbwt_router_host.hip (training and
// inference entry points) and kbwt_host.hip (features computed during the forward pass).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kbwt_internal.h"

namespace kbwt {

constexpr int ROUTER_NF = 262, ROUTER_NH = 64, ROUTER_NC = 4;
constexpr int ROUTER_NPARAM = ROUTER_NF * ROUTER_NH + ROUTER_NH + ROUTER_NH * ROUTER_NH + ROUTER_NH +
                              ROUTER_NH * ROUTER_NC + ROUTER_NC;   // 21,252
constexpr int ROUTER_ROUNDS = 5;                                     // m_r / n for r = 1..5
constexpr int BIGRAM_TILE = 65536;

void launch_count_active(const uint32_t *gh, const uint32_t *gs, uint32_t G, const uint16_t *blk, uint32_t *act,
                         hipStream_t s);
void launch_order1(const uint8_t *T, Tiling tl, uint32_t ntiles, const BlockInfo *blocks, uint32_t B, uint32_t *hist,
                   float *feat, hipStream_t s);
void launch_mlp(bool grad, const float *X, const int32_t *y, const float *wt, const uint32_t *idx, uint32_t ns,
                const float *P, float *G, float *probs, float *loss, hipStream_t s);
void launch_gate(const float *X, uint32_t ns, const float *protos, uint32_t np, float *dist, hipStream_t s);
void launch_reduce_adam(const float *G, uint32_t nb, float *P, float *M, float *V, float lr, float beta1, float beta2,
                        float eps, float bc1, float bc2, float *gout, hipStream_t s);

// Records a message for kbwt_last_error (kbwt_host.hip).
void record_error(const char *msg);

}  // namespace kbwt
