// Host side of the learned router: parameter initialisation, training (minibatch Adam on the GPU), the batch
// gradient and inference, behind the C-ABI of include/kbwt.h. Inputs and outputs are host arrays; every call
// allocates its device buffers, runs on the library's device and frees them (the data are small).
#include "kbwt_router.h"
#include "../../include/kbwt.h"

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

using namespace kbwt;

namespace {

struct RouterError {
    int code;
};

void check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    record_error(buf);
    throw RouterError{e == hipErrorOutOfMemory ? KBWT_ERR_NOMEM : KBWT_ERR_DEVICE};
}
#define RCHK(x) check((x), #x)

// Device buffer freed on scope exit.
struct DBuf {
    void *p = nullptr;
    explicit DBuf(size_t bytes) { RCHK(hipMalloc(&p, bytes < 4 ? 4 : bytes)); }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

template <class F> int32_t guarded_router(F f) {
    try {
        f();
        return 0;
    } catch (const RouterError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        record_error("host allocation failed");
        return KBWT_ERR_NOMEM;
    }
}

uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform in [-1, 1)
float unit(uint64_t &x) { return (float)((double)(splitmix(x) >> 11) * (2.0 / 9007199254740992.0) - 1.0); }

bool labels_ok(const int32_t *y, int32_t ns) {
    for (int32_t i = 0; i < ns; i++)
        if (y[i] < 0 || y[i] >= ROUTER_NC) return false;
    return true;
}

}  // namespace

extern "C" {

int32_t kbwt_router_init(float *params, uint64_t seed) {
    if (params == NULL) return KBWT_ERR_ARG;
    // He-uniform weights (ReLU layers), zero biases; layout as in kbwt_router.hip
    uint64_t x = seed;
    float *p = params;
    const int fan[3] = {ROUTER_NF, ROUTER_NH, ROUTER_NH}, out[3] = {ROUTER_NH, ROUTER_NH, ROUTER_NC};
    for (int l = 0; l < 3; l++) {
        const float a = std::sqrt(6.f / (float)fan[l]);
        for (int i = 0; i < fan[l] * out[l]; i++) *p++ = a * unit(x);
        for (int i = 0; i < out[l]; i++) *p++ = 0.f;
    }
    return 0;
}

int32_t kbwt_router_predict(const float *X, int32_t ns, const float *params, float *probs) {
    if (X == NULL || params == NULL || probs == NULL || ns < 0) return KBWT_ERR_ARG;
    if (ns == 0) return 0;
    return guarded_router([&] {
        DBuf dX((size_t)ns * ROUTER_NF * 4), dP((size_t)ROUTER_NPARAM * 4), dO((size_t)ns * ROUTER_NC * 4);
        RCHK(hipMemcpy(dX.p, X, (size_t)ns * ROUTER_NF * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dP.p, params, (size_t)ROUTER_NPARAM * 4, hipMemcpyHostToDevice));
        launch_mlp(false, dX.as<float>(), nullptr, nullptr, nullptr, (uint32_t)ns, dP.as<float>(), nullptr,
                   dO.as<float>(), nullptr, nullptr);
        RCHK(hipGetLastError());
        RCHK(hipMemcpy(probs, dO.p, (size_t)ns * ROUTER_NC * 4, hipMemcpyDeviceToHost));
    });
}

int32_t kbwt_router_grad(const float *X, const int32_t *y, const float *w, int32_t ns, const float *params,
                         float *grad, float *loss) {
    if (X == NULL || y == NULL || params == NULL || grad == NULL || ns <= 0 || ns > KBWT_ROUTER_GRAD_MAX ||
        !labels_ok(y, ns))
        return KBWT_ERR_ARG;
    return guarded_router([&] {
        DBuf dX((size_t)ns * ROUTER_NF * 4), dy((size_t)ns * 4), dw((size_t)ns * 4), dP((size_t)ROUTER_NPARAM * 4),
            dG((size_t)ns * ROUTER_NPARAM * 4), dg((size_t)ROUTER_NPARAM * 4), dl((size_t)ns * 4);
        RCHK(hipMemcpy(dX.p, X, (size_t)ns * ROUTER_NF * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dy.p, y, (size_t)ns * 4, hipMemcpyHostToDevice));
        if (w) RCHK(hipMemcpy(dw.p, w, (size_t)ns * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dP.p, params, (size_t)ROUTER_NPARAM * 4, hipMemcpyHostToDevice));
        launch_mlp(true, dX.as<float>(), dy.as<int32_t>(), w ? dw.as<float>() : nullptr, nullptr, (uint32_t)ns,
                   dP.as<float>(), dG.as<float>(), nullptr, dl.as<float>(), nullptr);
        launch_reduce_adam(dG.as<float>(), (uint32_t)ns, dP.as<float>(), nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f,
                           1.f, dg.as<float>(), nullptr);
        RCHK(hipGetLastError());
        RCHK(hipMemcpy(grad, dg.p, (size_t)ROUTER_NPARAM * 4, hipMemcpyDeviceToHost));
        if (loss) {
            std::vector<float> l(ns);
            RCHK(hipMemcpy(l.data(), dl.p, (size_t)ns * 4, hipMemcpyDeviceToHost));
            double sum = 0;
            for (float v : l) sum += v;
            *loss = (float)(sum / ns);
        }
    });
}

int32_t kbwt_router_gate(const float *X, int32_t ns, const float *protos, int32_t np, float *dist) {
    if (X == NULL || protos == NULL || dist == NULL || ns < 0 || np <= 0) return KBWT_ERR_ARG;
    if (ns == 0) return 0;
    return guarded_router([&] {
        DBuf dX((size_t)ns * ROUTER_NF * 4), dP((size_t)np * ROUTER_NF * 4), dD((size_t)ns * 4);
        RCHK(hipMemcpy(dX.p, X, (size_t)ns * ROUTER_NF * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dP.p, protos, (size_t)np * ROUTER_NF * 4, hipMemcpyHostToDevice));
        launch_gate(dX.as<float>(), (uint32_t)ns, dP.as<float>(), (uint32_t)np, dD.as<float>(), nullptr);
        RCHK(hipGetLastError());
        RCHK(hipMemcpy(dist, dD.p, (size_t)ns * 4, hipMemcpyDeviceToHost));
    });
}

int32_t kbwt_router_route(const uint8_t *const *blocks, const int32_t *n, int32_t count, const float *params,
                          float threshold, const float *protos, int32_t nprotos, float radius, int32_t *decision,
                          float *probs) {
    if (blocks == NULL || n == NULL || params == NULL || decision == NULL || count < 0 || count > 65535 ||
        (protos != NULL && nprotos <= 0))
        return KBWT_ERR_ARG;
    int64_t total = 0;
    for (int32_t b = 0; b < count; b++) {
        if (n[b] < 0 || (n[b] > 0 && blocks[b] == NULL)) return KBWT_ERR_ARG;
        total += n[b];
    }
    if (total >= ((int64_t)1 << 31) - 64) return KBWT_ERR_ARG;
    if (count == 0) return 0;
    return guarded_router([&] {
        // the frame's blocks back to back on the device, one sample each (r: the power of two at or above n)
        std::vector<kbwt_seg> segs(count);
        int64_t at = 0;
        for (int32_t b = 0; b < count; b++) {
            int32_t r = 2;
            while (r < n[b] && r < (1 << 30)) r <<= 1;
            segs[b] = {at, n[b], r};
            at += n[b];
        }
        DBuf dT((size_t)at + 64), dU((size_t)at + 64);
        for (int32_t b = 0; b < count; b++)
            if (n[b]) RCHK(hipMemcpy(dT.as<uint8_t>() + segs[b].offset, blocks[b], (size_t)n[b], hipMemcpyHostToDevice));
        std::vector<int32_t> I((size_t)count * 2);
        std::vector<float> F((size_t)count * ROUTER_NF), P((size_t)count * ROUTER_NC);
        int32_t rc = kbwt_dev_bwt_aux_batch_features(dT.as<uint8_t>(), dU.as<uint8_t>(), segs.data(), count, I.data(),
                                                     F.data(), nullptr);
        if (rc != 0) throw RouterError{rc};
        rc = kbwt_router_predict(F.data(), count, params, P.data());
        if (rc != 0) throw RouterError{rc};
        std::vector<float> D(protos ? count : 0);
        if (protos) {
            rc = kbwt_router_gate(F.data(), count, protos, nprotos, D.data());
            if (rc != 0) throw RouterError{rc};
        }
        for (int32_t b = 0; b < count; b++) {
            const float *p = P.data() + (size_t)b * ROUTER_NC;
            int32_t best = 0;
            for (int32_t c = 1; c < ROUTER_NC; c++)
                if (p[c] > p[best]) best = c;
            decision[b] = p[best] >= threshold && (!protos || D[b] <= radius) ? best : -1;
        }
        if (probs) memcpy(probs, P.data(), P.size() * 4);
    });
}

int32_t kbwt_router_train(const float *X, const int32_t *y, const float *w, int32_t ns,
                          const kbwt_router_train_cfg *cfg, float *params, float *epoch_loss) {
    if (X == NULL || y == NULL || cfg == NULL || params == NULL || ns <= 0 || cfg->epochs < 0 || cfg->batch <= 0 ||
        cfg->batch > KBWT_ROUTER_GRAD_MAX || !(cfg->lr > 0.f) || !labels_ok(y, ns))
        return KBWT_ERR_ARG;
    return guarded_router([&] {
        const uint32_t nb_max = (uint32_t)std::min(cfg->batch, ns);
        DBuf dX((size_t)ns * ROUTER_NF * 4), dy((size_t)ns * 4), dw((size_t)ns * 4), dP((size_t)ROUTER_NPARAM * 4),
            dM((size_t)ROUTER_NPARAM * 4), dV((size_t)ROUTER_NPARAM * 4), dG((size_t)nb_max * ROUTER_NPARAM * 4),
            dI((size_t)ns * 4), dl((size_t)ns * 4);
        RCHK(hipMemcpy(dX.p, X, (size_t)ns * ROUTER_NF * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dy.p, y, (size_t)ns * 4, hipMemcpyHostToDevice));
        if (w) RCHK(hipMemcpy(dw.p, w, (size_t)ns * 4, hipMemcpyHostToDevice));
        RCHK(hipMemcpy(dP.p, params, (size_t)ROUTER_NPARAM * 4, hipMemcpyHostToDevice));
        RCHK(hipMemset(dM.p, 0, (size_t)ROUTER_NPARAM * 4));
        RCHK(hipMemset(dV.p, 0, (size_t)ROUTER_NPARAM * 4));
        std::vector<uint32_t> order(ns);
        std::vector<float> l(ns);
        uint64_t rng = cfg->seed;
        const float b1 = cfg->beta1 > 0.f ? cfg->beta1 : 0.9f, b2 = cfg->beta2 > 0.f ? cfg->beta2 : 0.999f;
        const float eps = cfg->eps > 0.f ? cfg->eps : 1e-8f;
        double pb1 = 1.0, pb2 = 1.0;
        for (int32_t e = 0; e < cfg->epochs; e++) {
            for (int32_t i = 0; i < ns; i++) order[i] = (uint32_t)i;
            for (int32_t i = ns - 1; i > 0; i--) {   // Fisher-Yates with the seeded generator
                const uint32_t j = (uint32_t)(splitmix(rng) % (uint64_t)(i + 1));
                std::swap(order[i], order[j]);
            }
            RCHK(hipMemcpy(dI.p, order.data(), (size_t)ns * 4, hipMemcpyHostToDevice));
            for (int32_t b0 = 0; b0 < ns; b0 += (int32_t)nb_max) {
                const uint32_t nb = (uint32_t)std::min<int32_t>((int32_t)nb_max, ns - b0);
                launch_mlp(true, dX.as<float>(), dy.as<int32_t>(), w ? dw.as<float>() : nullptr, dI.as<uint32_t>() + b0,
                           nb, dP.as<float>(), dG.as<float>(), nullptr, dl.as<float>() + b0, nullptr);
                pb1 *= b1;
                pb2 *= b2;
                launch_reduce_adam(dG.as<float>(), nb, dP.as<float>(), dM.as<float>(), dV.as<float>(), cfg->lr, b1, b2,
                                   eps, (float)(1.0 - pb1), (float)(1.0 - pb2), nullptr, nullptr);
            }
            RCHK(hipGetLastError());
            if (epoch_loss) {
                RCHK(hipMemcpy(l.data(), dl.p, (size_t)ns * 4, hipMemcpyDeviceToHost));
                double sum = 0;
                for (float v : l) sum += v;
                epoch_loss[e] = (float)(sum / ns);
            }
        }
        RCHK(hipMemcpy(params, dP.p, (size_t)ROUTER_NPARAM * 4, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
