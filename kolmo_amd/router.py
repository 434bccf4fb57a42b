"""The learned router: which coder a level-8 block goes to, from features the GPU computes while it sorts the block.

kolmo decides per block between block sorting (codec 18) and KLZMA (codec 24) by coding a slice of the block both
ways (src/core/blocks.c:390-411, 473-507). Here a small fp32 MLP (262 -> 64 -> 64 -> 4), trained and run by the
library's HIP kernels (kbwt_router_*), reads per-block features that come out of the batched forward pass
(kbwt_dev_bwt_aux_batch_features) and predicts the class of the smallest output: 0 block sorting of the whole block,
1 KLZMA, 2 store, 3 block sorting at half the block size (the block's two halves as two blocks). Exhaustive
best-per-block, coding the block every way and keeping the smallest, stays the fallback: a block whose top probability
is below the threshold is coded every way, and so is a block the novelty gate finds unlike every training block (its
features farther than a radius from each of a few hundred prototypes picked from the training set; kbwt_router_gate).
The choice is recorded per block, so coders, containers and the decoder do not change.

Class labels come from running the coders (kolmo_bwt_compress with KOLMO_BWT_DEFAULTS on the block and on its halves,
kolmo_klzma_compress at level 8); the caller supplies them.
"""
import ctypes
import os

import numpy as np

from . import Seg, _check, _i32p, _ptr, lib, n_samples

NF, NC, NPARAM, GRAD_MAX = 262, 4, 21252, 4096
BWT, KLZMA, STORE, BWT2 = 0, 1, 2, 3
HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "data", "router_model.npz")
_f32p = ctypes.POINTER(ctypes.c_float)


class TrainCfg(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


_sigs_done = False


def _lib():
    global _sigs_done
    L = lib()
    if not _sigs_done:
        L.kbwt_dev_bwt_aux_batch_features.restype = ctypes.c_int32
        L.kbwt_dev_bwt_aux_batch_features.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Seg),
                                                      ctypes.c_int32, _i32p, _f32p, ctypes.c_void_p]
        L.kbwt_router_init.restype = ctypes.c_int32
        L.kbwt_router_init.argtypes = [_f32p, ctypes.c_uint64]
        L.kbwt_router_train.restype = ctypes.c_int32
        L.kbwt_router_train.argtypes = [_f32p, _i32p, _f32p, ctypes.c_int32, ctypes.POINTER(TrainCfg), _f32p, _f32p]
        L.kbwt_router_grad.restype = ctypes.c_int32
        L.kbwt_router_grad.argtypes = [_f32p, _i32p, _f32p, ctypes.c_int32, _f32p, _f32p, _f32p]
        L.kbwt_router_predict.restype = ctypes.c_int32
        L.kbwt_router_predict.argtypes = [_f32p, ctypes.c_int32, _f32p, _f32p]
        L.kbwt_router_route.restype = ctypes.c_int32
        L.kbwt_router_gate.restype = ctypes.c_int32
        L.kbwt_router_gate.argtypes = [_f32p, ctypes.c_int32, _f32p, ctypes.c_int32, _f32p]
        L.kbwt_router_route.argtypes = [ctypes.POINTER(ctypes.c_void_p), _i32p, ctypes.c_int32, _f32p, ctypes.c_float,
                                        _f32p, ctypes.c_int32, ctypes.c_float, _i32p, _f32p]
        _sigs_done = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def dev_features_batch(dT, segs, count, dU=None, stream=None):
    """Batched forward BWT over a device uint8 tensor and the router features of every block:
    returns (dU, I, features[count, 262])."""
    import torch
    if dU is None:
        dU = torch.empty_like(dT)
    I = np.zeros(max(sum(n_samples(segs[b].n, segs[b].r) for b in range(count)), 1), np.int32)
    F = np.zeros((max(count, 1), NF), np.float32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(_lib().kbwt_dev_bwt_aux_batch_features(dT.data_ptr(), dU.data_ptr(), segs, count, _ptr(I, _i32p),
                                                  _ptr(F, _f32p), s))
    return dU, I, F[:count]


def features_of(blocks, rs=None):
    """Features of host blocks (list of uint8 arrays), packed into one device batch."""
    import torch
    from . import codec18_rlog, make_segs
    sizes = [int(b.size) for b in blocks]
    rs = rs or [1 << codec18_rlog(max(n, 2)) for n in sizes]
    segs, total = make_segs(sizes, rs)
    host = np.concatenate([np.asarray(b, np.uint8) for b in blocks] + [np.zeros(1, np.uint8)])
    dT = torch.from_numpy(host).cuda()
    _, _, F = dev_features_batch(dT, segs, len(blocks))
    return F


def init(seed=1):
    P = np.zeros(NPARAM, np.float32)
    _check(_lib().kbwt_router_init(_ptr(P, _f32p), seed))
    return P


def train(X, y, w=None, params=None, epochs=60, batch=256, lr=2e-3, seed=7):
    """Minibatch Adam on the GPU; returns (params, per-epoch mean loss)."""
    X, y = _f32(X), np.ascontiguousarray(y, np.int32)
    P = _f32(init(seed) if params is None else params).copy()
    w = None if w is None else _f32(w)
    cfg = TrainCfg(epochs, batch, lr, 0.9, 0.999, 1e-8, seed)
    L = np.zeros(max(epochs, 1), np.float32)
    _check(_lib().kbwt_router_train(_ptr(X, _f32p), _ptr(y, _i32p), None if w is None else _ptr(w, _f32p), len(y),
                                    ctypes.byref(cfg), _ptr(P, _f32p), _ptr(L, _f32p)))
    return P, L[:epochs]


def grad(X, y, params, w=None):
    """Mean weighted cross-entropy and its gradient (GPU)."""
    X, y, P = _f32(X), np.ascontiguousarray(y, np.int32), _f32(params)
    w = None if w is None else _f32(w)
    g = np.zeros(NPARAM, np.float32)
    loss = ctypes.c_float(0)
    _check(_lib().kbwt_router_grad(_ptr(X, _f32p), _ptr(y, _i32p), None if w is None else _ptr(w, _f32p), len(y),
                                   _ptr(P, _f32p), _ptr(g, _f32p), ctypes.byref(loss)))
    return float(loss.value), g


def predict(X, params):
    X, P = _f32(X), _f32(params)
    out = np.zeros((max(len(X), 1), NC), np.float32)
    _check(_lib().kbwt_router_predict(_ptr(X, _f32p), len(X), _ptr(P, _f32p), _ptr(out, _f32p)))
    return out[: len(X)]


def gate(X, protos):
    """kbwt_router_gate: distance from each sample's features to the nearest prototype (GPU)."""
    X, Q = _f32(X).reshape(-1, NF), _f32(protos).reshape(-1, NF)
    d = np.zeros(max(len(X), 1), np.float32)
    _check(_lib().kbwt_router_gate(_ptr(X, _f32p), len(X), _ptr(Q, _f32p), len(Q), _ptr(d, _f32p)))
    return d[: len(X)]


def route(blocks, params, threshold, protos=None, radius=0.0):
    """kbwt_router_route: the decision of every host block (class, or -1: code it every way) and the probabilities;
    with protos, blocks farther than radius from every prototype get -1."""
    blocks = [np.ascontiguousarray(b, np.uint8) for b in blocks]
    ptrs = (ctypes.c_void_p * max(len(blocks), 1))(*[b.ctypes.data for b in blocks])
    n = np.array([b.size for b in blocks] or [0], np.int32)
    dec = np.zeros(max(len(blocks), 1), np.int32)
    probs = np.zeros((max(len(blocks), 1), NC), np.float32)
    P = _f32(params)
    Q = None if protos is None else _f32(protos).reshape(-1, NF)
    _check(_lib().kbwt_router_route(ptrs, _ptr(n, _i32p), len(blocks), _ptr(P, _f32p), float(threshold),
                                    None if Q is None else _ptr(Q, _f32p), 0 if Q is None else len(Q), float(radius),
                                    _ptr(dec, _i32p), _ptr(probs, _f32p)))
    return dec[: len(blocks)], probs[: len(blocks)]


# ---- the router's workload blocks

PAIR_BASE = 100   # kind PAIR_BASE + 16 a + b: the first half from generator a, the second from generator b


def workload_block(kind, seed, n, model_seed=5):
    """A block of the router's workload: generator `kind` (kolmo_amd.GEN_KINDS), or for kind >= PAIR_BASE two
    generators' blocks of half the size one after the other (the heterogeneous blocks where coding the halves apart
    can pay)."""
    from . import generate
    kind = int(kind)
    if kind < PAIR_BASE:
        return generate(kind, int(seed), int(n), model_seed=model_seed)
    a, b = divmod(kind - PAIR_BASE, 16)
    h = int(n) // 2
    return np.concatenate([generate(a, int(seed), h, model_seed=model_seed),
                           generate(b, int(seed) ^ 0x9E3779B97F4A7C15, int(n) - h, model_seed=model_seed)])


# ---- labels, weights and decisions (host arithmetic over coder sizes)

def outcome_sizes(n, size_bwt, size_klzma, size_bwt2):
    """Bytes each class costs per block (columns BWT, KLZMA, STORE, BWT2): a coder that could not shrink the block
    (negative size) stores it."""
    n = np.asarray(n, np.int64)
    sb = np.where(np.asarray(size_bwt) > 0, size_bwt, n)
    sk = np.where(np.asarray(size_klzma) > 0, size_klzma, n)
    s2 = np.where(np.asarray(size_bwt2) > 0, size_bwt2, n)
    return np.stack([sb, sk, n, s2], axis=1).astype(np.int64)


def labels(sizes):
    """Class of the smallest output; ties go to the cheaper action: storing, then block sorting of the whole block,
    then KLZMA, then block sorting of the halves (a block no coder can shrink is stored)."""
    order = np.array([STORE, BWT, KLZMA, BWT2])
    return order[np.argmin(sizes[:, order], axis=1)].astype(np.int32)


def weights(sizes):
    """Per-block loss weights: the bytes a wrong choice costs, relative to the block, mean 1 (so the training loss
    follows the size objective: close calls matter little, clear ones a lot)."""
    n = sizes[:, 2].astype(np.float64)
    regret = (sizes - sizes.min(axis=1, keepdims=True)).sum(axis=1) / (2.0 * n)
    w = 0.02 + regret
    return (w / w.mean()).astype(np.float32)


def decide(probs, threshold, dist=None, radius=None):
    """Argmax class, or -1 (code every way, keep the smallest) when the top probability is below threshold or, given
    the gate's distances, the block lies farther than radius from every prototype."""
    c = np.argmax(probs, axis=1)
    ok = probs.max(axis=1) >= threshold
    if dist is not None:
        ok &= np.asarray(dist) <= radius
    return np.where(ok, c, -1).astype(np.int32)


def prototypes(X, k, first=0):
    """k training samples picked by farthest-point sampling (each next one the sample farthest from those picked so
    far), and the coverage radius: the largest distance from any sample to its nearest prototype."""
    X = np.asarray(X, np.float64)
    pick = [int(first)]
    d = np.sqrt(((X - X[first]) ** 2).sum(1))
    for _ in range(1, min(k, len(X))):
        j = int(np.argmax(d))
        pick.append(j)
        d = np.minimum(d, np.sqrt(((X - X[j]) ** 2).sum(1)))
    return np.asarray(pick, np.int64), float(d.max())


def evaluate(decision, sizes, t_bwt, t_klzma, t_bwt2, t_features=None):
    """Total output bytes and coder seconds of a per-block decision (-1: every coder, the smallest kept)."""
    decision = np.asarray(decision)
    best = sizes.min(axis=1)
    pick = np.where(decision >= 0, sizes[np.arange(len(sizes)), np.maximum(decision, 0)], best)
    t = np.select([decision == BWT, decision == KLZMA, decision == STORE, decision == BWT2],
                  [t_bwt, t_klzma, 0.0, t_bwt2], t_bwt + t_klzma + t_bwt2)
    total_t = float(np.sum(t)) + (float(np.sum(t_features)) if t_features is not None else 0.0)
    return {"bytes": int(pick.sum()), "best_bytes": int(best.sum()),
            "loss_pct": float(100.0 * (pick.sum() - best.sum()) / best.sum()), "seconds": total_t,
            "fallback_blocks": int(np.sum(decision < 0))}


def save_model(path, params, threshold, meta, protos=None, radius=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    gate_ = {} if protos is None else {"protos": _f32(protos), "radius": np.float32(radius)}
    np.savez_compressed(path, params=_f32(params), threshold=np.float32(threshold), **gate_,
                        **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path=MODEL):
    d = np.load(path)
    return d["params"].astype(np.float32), float(d["threshold"])


def load_gate(path=MODEL):
    """The model's prototypes and radius, or (None, 0.0) for a model without a gate."""
    d = np.load(path)
    if "protos" not in d:
        return None, 0.0
    return d["protos"].astype(np.float32), float(d["radius"])
