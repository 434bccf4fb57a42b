"""The learned router: which coder a level-8 block goes to, from features the GPU computes while it sorts the block.

kolmo decides per block between block sorting (codec 18) and KLZMA (codec 24) by coding a slice of the block both
ways (src/core/blocks.c:390-411, 473-507). Here a small fp32 MLP (262 -> 64 -> 64 -> 3), trained and run by the
library's HIP kernels (kbwt_router_*), reads per-block features that come out of the batched forward pass
(kbwt_dev_bwt_aux_batch_features) and predicts the class of the smaller output: 0 block sorting, 1 KLZMA, 2 store.
Exhaustive best-per-block, coding the block both ways and keeping the smaller, stays the fallback: a block whose top
probability is below the threshold is coded both ways. The choice is recorded per block, so coders, containers and
the decoder do not change.

Class labels come from running both coders (kolmo_bwt_compress with KOLMO_BWT_DEFAULTS, kolmo_klzma_compress at level
8); the caller supplies them.
"""
import ctypes
import os

import numpy as np

from . import Seg, _check, _i32p, _ptr, lib, n_samples

NF, NC, NPARAM, GRAD_MAX = 262, 3, 21187, 4096
BWT, KLZMA, STORE = 0, 1, 2
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


# ---- labels, weights and decisions (host arithmetic over coder sizes)

def outcome_sizes(n, size_bwt, size_klzma):
    """Bytes each class costs per block: a coder that could not shrink the block (negative size) stores it."""
    n = np.asarray(n, np.int64)
    sb = np.where(np.asarray(size_bwt) > 0, size_bwt, n)
    sk = np.where(np.asarray(size_klzma) > 0, size_klzma, n)
    return np.stack([sb, sk, n], axis=1).astype(np.int64)


def labels(sizes):
    """Class of the smallest output; ties go to the cheaper action: storing, then block sorting, then KLZMA (a block
    neither coder can shrink is stored)."""
    order = np.array([STORE, BWT, KLZMA])
    return order[np.argmin(sizes[:, order], axis=1)].astype(np.int32)


def weights(sizes):
    """Per-block loss weights: the bytes a wrong choice costs, relative to the block, mean 1 (so the training loss
    follows the size objective: close calls matter little, clear ones a lot)."""
    n = sizes[:, 2].astype(np.float64)
    regret = (sizes - sizes.min(axis=1, keepdims=True)).sum(axis=1) / (2.0 * n)
    w = 0.02 + regret
    return (w / w.mean()).astype(np.float32)


def decide(probs, threshold):
    """Argmax class, or -1 (code both ways, keep the smaller) when the top probability is below threshold."""
    c = np.argmax(probs, axis=1)
    return np.where(probs.max(axis=1) >= threshold, c, -1).astype(np.int32)


def evaluate(decision, sizes, t_bwt, t_klzma, t_features=None):
    """Total output bytes and coder seconds of a per-block decision (-1: both coders, the smaller kept)."""
    decision = np.asarray(decision)
    best = sizes.min(axis=1)
    pick = np.where(decision >= 0, sizes[np.arange(len(sizes)), np.maximum(decision, 0)], best)
    t = np.where(decision == BWT, t_bwt, np.where(decision == KLZMA, t_klzma,
                                                  np.where(decision == STORE, 0.0, t_bwt + t_klzma)))
    total_t = float(np.sum(t)) + (float(np.sum(t_features)) if t_features is not None else 0.0)
    return {"bytes": int(pick.sum()), "best_bytes": int(best.sum()),
            "loss_pct": float(100.0 * (pick.sum() - best.sum()) / best.sum()), "seconds": total_t,
            "fallback_blocks": int(np.sum(decision < 0))}


def save_model(path, params, threshold, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=_f32(params), threshold=np.float32(threshold),
                        **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path=MODEL):
    d = np.load(path)
    return d["params"].astype(np.float32), float(d["threshold"])
