// The learned router: per-block features from the forward pass, and a small fp32 MLP (262 -> 64 -> 64 -> 3, ReLU)
// trained and run by the kernels below. It picks, per block, the coder a level-8 block goes to (block sorting,
// KLZMA, or storing the block), the choice kolmo's block pipeline makes today by coding slices both ways
// (src/core/blocks.c:390-411, 473-507); the choice is recorded per block, so neither coder nor the decoder changes.
//
// Features of block b (262 fp32):
//   [0, 256)   sqrt(count of byte c / n): the byte histogram of the forward's first kernel;
//   256        order-1 entropy H(T[i] | T[i-1]) / 8, from a bigram histogram;
//   [257, 262) m_r / n for r = 1..5: the share of suffixes still in a rank group at the start of doubling round r
//              (their first h0 * 2^(r-1) symbols repeat elsewhere in the block), counted from the round's group lists.
//
// The network is 21,187 parameters; one forward is ~42 kFLOP per block. Training is per-sample forward + backward
// (one wave per sample, gradients per sample) and a fixed-order batch reduction with the Adam update, so results do
// not depend on scheduling. At this size the work is launch-bound, not MFMA- or HBM-bound.
#include "kbwt_device.h"
#include "kbwt_internal.h"
#include "kbwt_router.h"

namespace kbwt {

// ------------------------------------------------------------------ features

// act[r * B + b] += rows of block b in the listed groups (heads gh, sizes gs); blk == nullptr: one block.
__global__ __launch_bounds__(256) void k_count_active(const uint32_t *__restrict__ gh, const uint32_t *__restrict__ gs,
                                                      uint32_t G, const uint16_t *__restrict__ blk,
                                                      uint32_t *__restrict__ act) {
    uint32_t cur = 0xFFFFFFFFu, sum = 0;
    for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < G; g += (size_t)gridDim.x * 256) {
        const uint32_t b = blk ? blk[gh[g]] : 0u;
        if (b != cur) {
            if (sum) atomicAdd(&act[cur], sum);
            cur = b;
            sum = 0;
        }
        sum += gs[g];
    }
    if (sum) atomicAdd(&act[cur], sum);
}
void launch_count_active(const uint32_t *gh, const uint32_t *gs, uint32_t G, const uint16_t *blk, uint32_t *act,
                         hipStream_t s) {
    uint32_t grid = (G + 255) / 256;
    if (grid > 256) grid = 256;
    if (grid) k_count_active<<<grid, 256, 0, s>>>(gh, gs, G, blk, act);
}

// Bigram histograms hist[b][y][x] = #{i in [1, n_b) : T[i-1] = y, T[i] = x}. Workgroup (tile, q) counts the pairs
// of its BIG_TILE positions whose previous byte lies in quarter q (64 values of y) in LDS, then adds the non-zero
// bins to the block's histogram.
constexpr int BIG_TILE = BIGRAM_TILE;
__global__ __launch_bounds__(256) void k_bigram_hist(const uint8_t *__restrict__ T, Tiling tl,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[64 * 256];
    const int tid = threadIdx.x;
    const uint32_t q = blockIdx.y;
    uint32_t lo = 0, hi = tl.B;   // the block of this tile
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tl.first[mid] <= blockIdx.x) lo = mid; else hi = mid;
    }
    const uint32_t b = lo;
    const BlockInfo bi = tl.blocks[b];
    const size_t beg = (size_t)(blockIdx.x - tl.first[b]) * BIG_TILE;
    const size_t end = beg + BIG_TILE < bi.n ? beg + BIG_TILE : bi.n;
    for (int i = tid; i < 64 * 256; i += 256) h[i] = 0;
    __syncthreads();
    const uint8_t *t = T + bi.off;
    for (size_t i = (beg ? beg : 1) + tid; i < end; i += 256) {
        const uint32_t y = t[i - 1];
        if ((y >> 6) == q) atomicAdd(&h[((y & 63u) << 8) | t[i]], 1u);
    }
    __syncthreads();
    uint32_t *dst = hist + (size_t)b * 65536 + ((size_t)q << 14);
    for (int i = tid; i < 64 * 256; i += 256)
        if (h[i]) atomicAdd(&dst[i], h[i]);
}

// feat[b * NF + 256] = H(x | y) / 8 from the bigram histogram, which is cleared for the next use.
__global__ __launch_bounds__(256) void k_order1_entropy(uint32_t *__restrict__ hist, const BlockInfo *__restrict__ blocks,
                                                        float *__restrict__ feat) {
    __shared__ float s_part[4];
    const uint32_t b = blockIdx.x, y = threadIdx.x;
    uint32_t *row = hist + (size_t)b * 65536 + (size_t)y * 256;
    uint32_t cy = 0;
    for (int x = 0; x < 256; x++) cy += row[x];
    float acc = 0.f;
    if (cy) {
        const float lcy = __log2f((float)cy);
        for (int x = 0; x < 256; x++) {
            const uint32_t c = row[x];
            if (c) acc += (float)c * (lcy - __log2f((float)c));
            row[x] = 0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
    if ((y & 63) == 0) s_part[y >> 6] = acc;
    __syncthreads();
    if (y == 0) {
        const uint32_t pairs = blocks[b].n > 1 ? blocks[b].n - 1 : 1;
        feat[(size_t)b * ROUTER_NF + 256] = (s_part[0] + s_part[1] + s_part[2] + s_part[3]) / (float)pairs / 8.f;
    }
}
void launch_order1(const uint8_t *T, Tiling tl, uint32_t ntiles, const BlockInfo *blocks, uint32_t B, uint32_t *hist,
                   float *feat, hipStream_t s) {
    if (ntiles) k_bigram_hist<<<dim3(ntiles, 4), 256, 0, s>>>(T, tl, hist);
    if (B) k_order1_entropy<<<B, 256, 0, s>>>(hist, blocks, feat);
}

// ------------------------------------------------------------------ the network

// Parameters, input-major: W1[k * NH + j] (k < NF), b1, W2[k * NH + j], b2, W3[k * NC + c], b3.
constexpr int NF = ROUTER_NF, NH = ROUTER_NH, NC = ROUTER_NC;
constexpr int OW1 = 0, OB1 = OW1 + NF * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH * NC;
static_assert(OB3 + NC == ROUTER_NPARAM, "parameter layout");

// One wave per sample s (row idx[s] of X, or s): softmax probabilities, the weighted cross-entropy wt * -log p[y],
// and with GRAD the sample's gradient of that loss into G[s * NPARAM ..].
template <bool GRAD>
__global__ __launch_bounds__(64) void k_mlp(const float *__restrict__ X, const int32_t *__restrict__ y,
                                            const float *__restrict__ wt, const uint32_t *__restrict__ idx,
                                            uint32_t ns, const float *__restrict__ P, float *__restrict__ G,
                                            float *__restrict__ probs, float *__restrict__ loss) {
    __shared__ float x[NF], h1[NH], h2[NH], d2[NH], dz[NC], pz[NC];
    const uint32_t s = blockIdx.x;
    if (s >= ns) return;
    const uint32_t row = idx ? idx[s] : s;
    const int j = threadIdx.x;
    for (int k = j; k < NF; k += NH) x[k] = X[(size_t)row * NF + k];
    __syncthreads();
    float a1 = P[OB1 + j];
    for (int k = 0; k < NF; k++) a1 = fmaf(P[OW1 + k * NH + j], x[k], a1);
    h1[j] = a1 > 0.f ? a1 : 0.f;
    __syncthreads();
    float a2 = P[OB2 + j];
    for (int k = 0; k < NH; k++) a2 = fmaf(P[OW2 + k * NH + j], h1[k], a2);
    h2[j] = a2 > 0.f ? a2 : 0.f;
    __syncthreads();
    if (j < NC) {
        float z = P[OB3 + j];
        for (int k = 0; k < NH; k++) z = fmaf(P[OW3 + k * NC + j], h2[k], z);
        pz[j] = z;
    }
    __syncthreads();
    if (j == 0) {
        float m = pz[0];
        for (int c = 1; c < NC; c++) m = fmaxf(m, pz[c]);
        float p[NC], se = 0.f;
        for (int c = 0; c < NC; c++) {
            p[c] = __expf(pz[c] - m);
            se += p[c];
        }
        for (int c = 0; c < NC; c++) p[c] /= se;
        const float w = wt ? wt[row] : 1.f;
        const int lab = y ? y[row] : 0;
        if (probs)
            for (int c = 0; c < NC; c++) probs[(size_t)s * NC + c] = p[c];
        if (loss) loss[s] = w * (m + __logf(se) - pz[lab]);
        for (int c = 0; c < NC; c++) dz[c] = w * (p[c] - (c == lab ? 1.f : 0.f));
    }
    if (!GRAD) return;
    __syncthreads();
    float *g = G + (size_t)s * ROUTER_NPARAM;
    // layer 3
    for (int c = 0; c < NC; c++) g[OW3 + j * NC + c] = h2[j] * dz[c];
    if (j < NC) g[OB3 + j] = dz[j];
    float dh2 = 0.f;
    if (a2 > 0.f)
        for (int c = 0; c < NC; c++) dh2 = fmaf(P[OW3 + j * NC + c], dz[c], dh2);
    d2[j] = dh2;
    __syncthreads();
    // layer 2 (thread j writes column j of every input row: coalesced)
    for (int k = 0; k < NH; k++) g[OW2 + k * NH + j] = h1[k] * d2[j];
    g[OB2 + j] = d2[j];
    float dh1 = 0.f;
    if (a1 > 0.f)
        for (int m = 0; m < NH; m++) dh1 = fmaf(P[OW2 + j * NH + m], d2[m], dh1);
    // layer 1
    for (int k = 0; k < NF; k++) g[OW1 + k * NH + j] = x[k] * dh1;
    g[OB1 + j] = dh1;
}

// Mean of the nb per-sample gradients in sample order; with M the Adam step (bias corrections bc1, bc2), with gout
// the mean gradient itself.
__global__ __launch_bounds__(256) void k_reduce_adam(const float *__restrict__ G, uint32_t nb, float *__restrict__ P,
                                                     float *__restrict__ M, float *__restrict__ V, float lr,
                                                     float beta1, float beta2, float eps, float bc1, float bc2,
                                                     float *__restrict__ gout) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= (uint32_t)ROUTER_NPARAM) return;
    float g = 0.f;
    for (uint32_t s = 0; s < nb; s++) g += G[(size_t)s * ROUTER_NPARAM + p];
    g /= (float)nb;
    if (gout) gout[p] = g;
    if (M) {
        const float m = beta1 * M[p] + (1.f - beta1) * g;
        const float v = beta2 * V[p] + (1.f - beta2) * g * g;
        M[p] = m;
        V[p] = v;
        P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    }
}

// The novelty gate: dist[s] = the Euclidean distance from sample s to the nearest of np prototype feature vectors
// (protos[p * NF ..]). One wave per sample; lane l sums the squared differences of features l, l + 64, ...; the wave
// adds them up with shuffles. The byte-histogram features are square roots of byte shares, so over them the distance
// is sqrt(2) times the Hellinger distance of the byte distributions.
__global__ __launch_bounds__(64) void k_gate(const float *__restrict__ X, uint32_t ns, const float *__restrict__ protos,
                                             uint32_t np, float *__restrict__ dist) {
    const uint32_t s = blockIdx.x;
    if (s >= ns) return;
    const int lane = threadIdx.x;
    constexpr int PER = (NF + 63) / 64;
    float x[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) x[q] = lane + 64 * q < NF ? X[(size_t)s * NF + lane + 64 * q] : 0.f;
    float best = 3.4e38f;
    for (uint32_t p = 0; p < np; p++) {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const float d = lane + 64 * q < NF ? x[q] - protos[(size_t)p * NF + lane + 64 * q] : 0.f;
            acc = fmaf(d, d, acc);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        best = fminf(best, acc);
    }
    if (lane == 0) dist[s] = sqrtf(best);
}

void launch_gate(const float *X, uint32_t ns, const float *protos, uint32_t np, float *dist, hipStream_t s) {
    if (ns) k_gate<<<ns, 64, 0, s>>>(X, ns, protos, np, dist);
}

void launch_mlp(bool grad, const float *X, const int32_t *y, const float *wt, const uint32_t *idx, uint32_t ns,
                const float *P, float *G, float *probs, float *loss, hipStream_t s) {
    if (!ns) return;
    if (grad) k_mlp<true><<<ns, NH, 0, s>>>(X, y, wt, idx, ns, P, G, probs, loss);
    else k_mlp<false><<<ns, NH, 0, s>>>(X, y, wt, idx, ns, P, G, probs, loss);
}
void launch_reduce_adam(const float *G, uint32_t nb, float *P, float *M, float *V, float lr, float beta1, float beta2,
                        float eps, float bc1, float bc2, float *gout, hipStream_t s) {
    k_reduce_adam<<<(ROUTER_NPARAM + 255) / 256, 256, 0, s>>>(G, nb, P, M, V, lr, beta1, beta2, eps, bc1, bc2, gout);
}

}  // namespace kbwt
