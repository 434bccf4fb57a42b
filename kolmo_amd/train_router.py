#!/usr/bin/env python3
"""Train the learned router on the GPU and write the model the library ships (kolmo_amd/data/router_model.npz).

  python -m kolmo_amd.train_router --labels tests/golden/router_labels.npz \\
      [--features FILE] [--features-out FILE] [--epochs 80]

Steps:
  1. features of every labelled block: the blocks are regenerated from their (generator, seed, size) and sorted in
     device batches by kbwt_dev_bwt_aux_batch_features (or read from --features, as written by --features-out);
  2. an 80/20 split by block index (every fifth block is held out);
  3. minibatch Adam on the weighted cross-entropy of the training blocks (kbwt_router_train);
  4. the fallback threshold: the largest top-probability threshold under which coding the unsure training blocks
     both ways keeps the training set within 0.25% of exhaustive best-per-block, preferring the fewest fallbacks;
  5. the novelty gate: --protos training blocks picked by farthest-point sampling, and a radius of --radius-factor
     times their coverage (the largest distance from a training block to its nearest prototype); a block farther
     than the radius from every prototype is coded every way (kbwt_router_gate);
  6. the held-out blocks: bytes and coder time against exhaustive best-per-block and against level 8's rule;
  7. with --ood, the same on blocks none of the generators made (real files of kolmo's tests, with their features);
  8. with --extra, more labelled blocks with their features (the heterogeneous pairs, make_router_labels.py --pairs)
     join the training data; their own held-out fifth is reported apart, and the 8192 generated blocks' held-out
     fifth stays the headline evaluation.
`--add-features LABELS OUT` computes the features of a labels file's blocks on the GPU and writes LABELS plus
`features` and `gpu_seconds` to OUT.
The labels' coder times are single-thread CPU times of kolmo's own coders, measured when the labels were made.
Classes: block sorting of the whole block, KLZMA, store, block sorting of the two halves (kolmo_amd.router).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import kolmo_amd as K  # noqa: E402
from kolmo_amd import router as R  # noqa: E402

ROUTER_SEED = 5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def compute_features(lab, chunk_bytes=768 << 20):
    import torch
    n = lab["n"].astype(np.int64)
    F = np.zeros((len(n), R.NF), np.float32)
    i = 0
    t_gpu = 0.0
    while i < len(n):
        j = i
        tot = 0
        while j < len(n) and (j == i or tot + n[j] <= chunk_bytes):
            tot += int(n[j])
            j += 1
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=16) as ex:
            blocks = list(ex.map(lambda b: R.workload_block(lab["kind"][b], lab["seed"][b], n[b],
                                                            model_seed=ROUTER_SEED), range(i, j)))
        sizes = [int(b.size) for b in blocks]
        rs = [1 << K.codec18_rlog(s) for s in sizes]
        segs, total = K.make_segs(sizes, rs)
        dT = torch.from_numpy(np.concatenate(blocks + [np.zeros(1, np.uint8)])).cuda()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, Fc = R.dev_features_batch(dT, segs, len(blocks))
        torch.cuda.synchronize()
        t_gpu += time.perf_counter() - t0
        F[i:j] = Fc
        log(f"features: blocks {i}-{j} of {len(n)} ({tot / 2**20:.0f} MiB), GPU {t_gpu:.2f} s so far")
        del dT
        i = j
    return F, t_gpu


def level8_rule(lab, sizes):
    """src/core/blocks.c:390-402: text goes to block sorting; another block of 64 KiB or more keeps block sorting only
    when its slice codes at least 2% smaller than with KLZMA (role level 1); smaller blocks go to block sorting.
    Returns the decision and the coder seconds (the slice trials included)."""
    n = lab["n"]
    text = lab["text"].astype(bool)
    sb, sk = lab["slice_bwt"].astype(np.int64), lab["slice_klzma1"].astype(np.int64)
    trial = ~text & (n >= (64 << 10))
    sbx = np.where(sb > 0, sb, n)
    skx = np.where(sk > 0, sk, n)
    lz = trial & (sbx * 100 > skx * 98)
    dec = np.where(lz, R.KLZMA, R.BWT)
    t = np.where(trial, lab["t_slice_bwt"] + lab["t_slice_klzma1"], 0.0)
    t = t + np.where(dec == R.KLZMA, lab["t_klzma"], lab["t_bwt"])
    pick = sizes[np.arange(len(n)), dec]
    return dec, float(t.sum()), int(pick.sum())


def choose_threshold(probs, sizes, t_b, t_k, t_2, budget_pct=0.25):
    best = None
    for th in np.linspace(0.25, 0.995, 150):
        ev = R.evaluate(R.decide(probs, th), sizes, t_b, t_k, t_2)
        if ev["loss_pct"] <= budget_pct and (best is None or ev["seconds"] < best[1]["seconds"]):
            best = (float(th), ev)
    return best if best is not None else (1.01, R.evaluate(R.decide(probs, 1.01), sizes, t_b, t_k, t_2))


def ood_report(path, P, th, protos, radius):
    """The model on the out-of-distribution blocks (tests/golden/router_ood.npz, with features): bytes and coder
    seconds against exhaustive best-per-block and always block sorting, with the gate and without it, and the
    accuracy of the network's argmax."""
    d = np.load(path)
    sizes = R.outcome_sizes(d["n"], d["size_bwt"], d["size_klzma"], d["size_bwt2"])
    y = R.labels(sizes)
    tb, tk, t2 = (d[k].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2"))
    probs = R.predict(d["features"], P)
    dist = R.gate(d["features"], protos)
    dec = R.decide(probs, th, dist, radius)
    raw = R.decide(probs, th)
    ev = R.evaluate(dec, sizes, tb, tk, t2)
    ex = R.evaluate(np.full(len(y), -1), sizes, tb, tk, t2)
    bwt = R.evaluate(np.full(len(y), R.BWT), sizes, tb, tk, t2)
    return {"blocks": int(len(y)), "bytes_in": int(d["n"].sum()), "labels": np.bincount(y, minlength=R.NC).tolist(),
            "accuracy": float(np.mean(np.argmax(probs, 1) == y)), "router": ev, "exhaustive": ex,
            "always_bwt": bwt, "router_over_exhaustive_throughput": ex["seconds"] / ev["seconds"],
            "without_gate": R.evaluate(raw, sizes, tb, tk, t2), "gate_distance_min": float(dist.min()),
            "per_block": [{"name": str(nm), "label": int(a), "decision": int(b), "argmax": int(c),
                           "gate_distance": round(float(g), 4)}
                          for nm, a, b, c, g in zip(d["name"], y, dec, np.argmax(probs, 1), dist)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--labels", default=os.path.join(os.path.dirname(HERE), "tests", "golden", "router_labels.npz"))
    ap.add_argument("--features")
    ap.add_argument("--features-out")
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--model-out", default=R.MODEL)
    ap.add_argument("--report")
    ap.add_argument("--ood", help="out-of-distribution blocks with labels and features (make_router_ood.py)")
    ap.add_argument("--extra", help="more labelled training blocks with their features")
    ap.add_argument("--add-features", nargs=2, metavar=("LABELS", "OUT"))
    ap.add_argument("--protos", type=int, default=256)
    ap.add_argument("--radius-factor", type=float, default=2.0)
    a = ap.parse_args()
    if a.add_features:
        d = dict(np.load(a.add_features[0]))
        F, t = compute_features(d)
        d.update(features=F, gpu_seconds=np.float64(t))
        np.savez_compressed(a.add_features[1], **d)
        log(f"features of {len(F)} blocks in {t:.2f} s of GPU time -> {a.add_features[1]}")
        return
    lab = dict(np.load(a.labels))
    nb = len(lab["n"])
    if a.features:
        F = np.load(a.features)["features"].astype(np.float32)
        t_feat = float(np.load(a.features)["gpu_seconds"])
    else:
        F, t_feat = compute_features(lab)
        if a.features_out:
            np.savez_compressed(a.features_out, features=F.astype(np.float32), gpu_seconds=np.float64(t_feat))
    sizes = R.outcome_sizes(lab["n"], lab["size_bwt"], lab["size_klzma"], lab["size_bwt2"])
    y = R.labels(sizes)
    w = R.weights(sizes)
    test = np.arange(nb) % 5 == 4
    tr = ~test
    Xtr, ytr, wtr = F[tr], y[tr], w[tr]
    if a.extra:
        xd = np.load(a.extra)
        xs = R.outcome_sizes(xd["n"], xd["size_bwt"], xd["size_klzma"], xd["size_bwt2"])
        xtest = np.arange(len(xs)) % 5 == 4
        Xtr = np.concatenate([Xtr, xd["features"][~xtest]])
        ytr = np.concatenate([ytr, R.labels(xs)[~xtest]])
        wtr = np.concatenate([wtr, R.weights(xs)[~xtest]])
    t0 = time.perf_counter()
    P, L = R.train(Xtr, ytr, wtr, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    t_train = time.perf_counter() - t0
    log(f"trained {a.epochs} epochs on {len(ytr)} blocks in {t_train:.1f} s; loss {L[0]:.4f} -> {L[-1]:.4f}")
    tb, tk, t2 = (lab[k].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2"))
    th, ev_tr = choose_threshold(R.predict(F[tr], P), sizes[tr], tb[tr], tk[tr], t2[tr])
    pick, cover = R.prototypes(Xtr, a.protos)
    protos, radius = Xtr[pick].astype(np.float32), a.radius_factor * cover
    log(f"gate: {len(pick)} prototypes cover the training blocks within {cover:.4f}; radius {radius:.4f}")
    probs = R.predict(F[test], P)
    dist = R.gate(F[test], protos)
    dec = R.decide(probs, th, dist, radius)
    share_feat = t_feat * (lab["n"][test].sum() / lab["n"].sum())
    ev = R.evaluate(dec, sizes[test], tb[test], tk[test], t2[test], t_features=[share_feat])
    ex = R.evaluate(np.full(test.sum(), -1), sizes[test], tb[test], tk[test], t2[test])
    rep = {"blocks": int(nb), "train_blocks": int(len(ytr)), "test_blocks": int(test.sum()),
           "test_bytes_in": int(lab["n"][test].sum()), "threshold": th, "epochs": a.epochs,
           "train_loss_first_last": [float(L[0]), float(L[-1])],
           "accuracy_test": float(np.mean(np.argmax(probs, 1) == y[test])),
           "labels_test": np.bincount(y[test], minlength=R.NC).tolist(),
           "router": ev, "exhaustive": ex,
           "router_over_exhaustive_throughput": ex["seconds"] / ev["seconds"],
           "features_gpu_seconds_test_share": share_feat,
           "gate": {"prototypes": int(len(pick)), "coverage": cover, "radius": radius,
                    "test_distance_max": float(dist.max()), "test_gated": int(np.sum(dist > radius))}}
    if "text" in lab:
        d8, t8, b8 = level8_rule({k: v[test] for k, v in lab.items()}, sizes[test])
        rep["level8_rule"] = {"bytes": b8, "loss_pct": 100.0 * (b8 - ex["best_bytes"]) / ex["best_bytes"],
                              "seconds": t8}
    rep["always_bwt"] = R.evaluate(np.full(test.sum(), R.BWT), sizes[test], tb[test], tk[test], t2[test])
    if a.extra:
        xt = [xd[k][xtest].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2")]
        xp = R.predict(xd["features"][xtest], P)
        xg = R.gate(xd["features"][xtest], protos)
        xe = R.evaluate(R.decide(xp, th, xg, radius), xs[xtest], *xt)
        xx = R.evaluate(np.full(xtest.sum(), -1), xs[xtest], *xt)
        rep["pairs_held_out"] = {"blocks": int(xtest.sum()), "labels": np.bincount(R.labels(xs)[xtest], minlength=R.NC).tolist(),
                                 "accuracy": float(np.mean(np.argmax(xp, 1) == R.labels(xs)[xtest])),
                                 "router": xe, "exhaustive": xx,
                                 "always_bwt": R.evaluate(np.full(xtest.sum(), R.BWT), xs[xtest], *xt),
                                 "router_over_exhaustive_throughput": xx["seconds"] / xe["seconds"],
                                 "gate_distance_max": float(xg.max())}
    if a.ood:
        rep["out_of_distribution"] = ood_report(a.ood, P, th, protos, radius)
    R.save_model(a.model_out, P, th, {"epochs": a.epochs, "lr": a.lr, "batch": a.batch, "seed": a.seed},
                 protos=protos, radius=radius)
    out = json.dumps(rep, indent=1)
    print(out)
    if a.report:
        open(a.report, "w").write(out + "\n")


if __name__ == "__main__":
    main()
