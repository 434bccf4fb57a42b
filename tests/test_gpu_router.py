"""The learned router on the GPU against its CPU oracle (oracle/router_oracle.py, float64), and what the trained
network gains.

- gradient of the weighted cross-entropy (kbwt_router_grad) against back-propagation in float64: max abs error
  <= 2e-5 * max |g| + 1e-7 (fp32 accumulation over 21,187 parameters and a few hundred samples);
- probabilities (kbwt_router_predict): abs error <= 2e-6;
- features (kbwt_dev_bwt_aux_batch_features): byte histogram within 1e-6, order-1 entropy within 2e-5 (fp32
  logarithms), m_r / n exactly; a block's features do not depend on the rest of its batch; the BWT of the call is the
  plain batch's;
- training lowers the loss and learns a rule the features carry;
- on the held-out fifth of the router workload (committed labels and features), the shipped model stays within 0.5%
  of exhaustive best-per-block bytes at >= 3x its coder throughput, and beats level 8's slice rule.
"""
import os
import sys

import numpy as np
import pytest

import kolmo_amd as K
from kolmo_amd import router as R

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import router_oracle as RO  # noqa: E402
import refs  # noqa: E402

pytestmark = pytest.mark.gpu
LABELS = os.path.join(HERE, "golden", "router_labels.npz")
FEATURES = os.path.join(HERE, "golden", "router_features.npz")


def say(*a):
    print(*a, flush=True)


def problem(seed, ns):
    rng = np.random.default_rng(seed)
    P = R.init(seed) + rng.normal(0, 0.02, R.NPARAM).astype(np.float32)
    X = rng.random((ns, R.NF)).astype(np.float32)
    X[:, :256] *= 0.2
    y = rng.integers(0, 3, ns).astype(np.int32)
    w = (rng.random(ns) + 0.5).astype(np.float32)
    return P, X, y, w


def test_gpu_router_gradient(gpu):
    for seed, ns in [(1, 1), (2, 37), (3, 300), (4, 4096)]:
        P, X, y, w = problem(seed, ns)
        say("router gradient", ns)
        L, g = R.grad(X, y, P, w)
        eL, eg = RO.loss_grad(P.astype(np.float64), X.astype(np.float64), y, w.astype(np.float64))
        assert abs(L - eL) <= 1e-5 * abs(eL) + 1e-6
        err = np.max(np.abs(g - eg))
        assert err <= 2e-5 * np.max(np.abs(eg)) + 1e-7, (ns, err, np.max(np.abs(eg)))


def test_gpu_router_predict(gpu):
    P, X, _, _ = problem(5, 513)
    p = R.predict(X, P)
    ep, _ = RO.forward(P.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(p - ep)) <= 2e-6
    assert np.allclose(p.sum(1), 1, atol=1e-6)


def feature_blocks():
    blocks = []
    for i, (kind, n) in enumerate([("markov2", 200000), ("csvcols", 90000), ("logs", 120000), ("bf16", 65536),
                                   ("random", 70000), ("dna", 150000), ("repeats", 180000), ("mixed", 100000),
                                   ("markov2", 5000), ("dna", 4096)]):
        blocks.append(K.generate(kind, 40 + i, n, model_seed=5))
    return blocks


def test_gpu_router_features(gpu):
    blocks = feature_blocks()
    say("router features of", len(blocks), "blocks")
    F = R.features_of(blocks)
    for b, T in enumerate(blocks):
        e = RO.features(T)
        assert np.max(np.abs(F[b, :256] - e[:256])) <= 1e-6, b
        assert abs(F[b, 256] - e[256]) <= 2e-5, (b, F[b, 256], e[256])
        got = np.round(F[b, 257:].astype(np.float64) * T.size)   # fp32 shares back to counts
        assert np.array_equal(got, np.round(e[257:] * T.size)), (b, got, e[257:] * T.size)


def test_gpu_router_features_batch_independent(gpu):
    blocks = feature_blocks()
    F = R.features_of(blocks)
    for b in [0, 5, 6]:
        F1 = R.features_of([blocks[b]])
        assert np.array_equal(F1[0], F[b]), b
    # the BWT that comes with the features is the plain batch's
    import torch
    sizes = [int(x.size) for x in blocks]
    rs = [1 << K.codec18_rlog(n) for n in sizes]
    segs, total = K.make_segs(sizes, rs)
    dT = torch.from_numpy(np.concatenate(blocks + [np.zeros(1, np.uint8)])).cuda()
    dU1, I1, _ = R.dev_features_batch(dT, segs, len(blocks))
    dU2, I2 = K.dev_bwt_aux_batch(dT, segs, len(blocks))
    torch.cuda.synchronize()
    assert torch.equal(dU1[:total], dU2[:total]) and np.array_equal(I1, I2)


def test_gpu_router_learns(gpu):
    """Labels from a rule the features carry (three thresholds over entropy, the repeat share and one more feature,
    four classes): the loss falls and the rule is learnt. 9,000 training samples: with 2,400 the network fits the
    noise in the byte-histogram features (held-out accuracy near 0.88)."""
    rng = np.random.default_rng(9)
    ns = 9600
    X = np.zeros((ns, R.NF), np.float32)
    X[:, :256] = rng.random((ns, 256)) * 0.1
    X[:, 256] = rng.random(ns)
    X[:, 257:] = rng.random((ns, 5))
    y = np.where(X[:, 256] > 0.9, R.STORE, np.where(X[:, 257] > 0.5, np.where(X[:, 258] > 0.7, R.BWT2, R.BWT),
                                                    R.KLZMA)).astype(np.int32)
    P, L = R.train(X[:9000], y[:9000], epochs=30, batch=128, lr=3e-3, seed=3)
    assert L[-1] < 0.5 * L[0]
    acc = np.mean(np.argmax(R.predict(X[9000:], P), 1) == y[9000:])
    say("router learns: loss", L[0], "->", L[-1], "held-out accuracy", acc)
    assert acc > 0.92


def test_gpu_router_gate_distances(gpu):
    """kbwt_router_gate against brute force on the host: the distance to the nearest prototype, for one prototype,
    for prototypes equal to samples (distance 0) and for more samples than a launch's blocks."""
    rng = np.random.default_rng(11)
    X = rng.random((700, R.NF)).astype(np.float32)
    Q = rng.random((37, R.NF)).astype(np.float32)
    Q[5] = X[3]
    want = np.sqrt(((X[:, None, :].astype(np.float64) - Q[None].astype(np.float64)) ** 2).sum(-1)).min(1)
    got = R.gate(X, Q)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5) and got[3] <= 1e-3
    one = R.gate(X[:9], Q[:1])
    assert np.allclose(one, np.sqrt(((X[:9] - Q[0]) ** 2).sum(1)), rtol=1e-5, atol=1e-5)
    assert R.gate(X[:0], Q).shape == (0,)


@pytest.mark.skipif(not (os.path.exists(LABELS) and os.path.exists(FEATURES) and os.path.exists(R.MODEL)),
                    reason="router labels, features or model not present")
def test_gpu_router_gain(gpu):
    lab = dict(np.load(LABELS))
    F = np.load(FEATURES)["features"].astype(np.float32)
    t_feat = float(np.load(FEATURES)["gpu_seconds"])
    P, th = R.load_model()
    sizes = R.outcome_sizes(lab["n"], lab["size_bwt"], lab["size_klzma"], lab["size_bwt2"])
    test = np.arange(len(F)) % 5 == 4
    tb, tk, t2 = (lab[k].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2"))
    # features of a sample of held-out blocks recomputed now equal the committed ones
    idx = np.flatnonzero(test)[:24]
    blocks = [R.workload_block(lab["kind"][i], lab["seed"][i], lab["n"][i], model_seed=5) for i in idx]
    Fs = R.features_of(blocks)
    assert np.max(np.abs(Fs - F[idx])) <= 1e-6
    protos, radius = R.load_gate()
    dec = R.decide(R.predict(F[test], P), th, R.gate(F[test], protos), radius)
    share = t_feat * lab["n"][test].sum() / lab["n"].sum()
    ev = R.evaluate(dec, sizes[test], tb[test], tk[test], t2[test], t_features=[share])
    ex = R.evaluate(np.full(test.sum(), -1), sizes[test], tb[test], tk[test], t2[test])
    bwt = R.evaluate(np.full(test.sum(), R.BWT), sizes[test], tb[test], tk[test], t2[test])
    say("router:", ev, "exhaustive:", ex, "always block sorting:", bwt)
    assert ev["loss_pct"] <= 0.5
    assert ex["seconds"] / ev["seconds"] >= 3.0
    assert ev["bytes"] < bwt["bytes"]
    if "text" in lab:
        from kolmo_amd.train_router import level8_rule
        _, t8, b8 = level8_rule({k: v[test] for k, v in lab.items()}, sizes[test])
        say("level 8 rule: bytes", b8, "seconds", t8)
        assert ev["bytes"] <= b8 and ev["seconds"] < t8


OOD = os.path.join(HERE, "golden", "router_ood.npz")


@pytest.mark.skipif(not (os.path.exists(OOD) and refs.have_full("ref") and os.path.exists(R.MODEL)),
                    reason="out-of-distribution set, kolmo's library or the model not present")
def test_gpu_router_out_of_distribution(gpu):
    """Real files of kolmo's own tests (tests/golden/make_router_ood.py): their features recomputed now equal the
    committed ones, and the shipped model's bytes and coder time on them are the ones the training report records."""
    import json
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_router_ood
    d = np.load(OOD)
    blocks = make_router_ood.ood_blocks()
    assert [b[0] for b in blocks] == list(d["name"])
    F = R.features_of([b[1] for b in blocks])
    assert np.max(np.abs(F - d["features"])) <= 1e-6
    P, th = R.load_model()
    protos, radius = R.load_gate()
    from kolmo_amd.train_router import ood_report
    rep = ood_report(OOD, P, th, protos, radius)
    say("out of distribution:", {k: rep[k] for k in ("blocks", "labels", "accuracy")}, rep["router"], rep["exhaustive"])
    want = json.load(open(os.path.join(HERE, "golden", "router_report.json")))["out_of_distribution"]
    assert rep["router"]["bytes"] == want["router"]["bytes"]
    assert rep["exhaustive"]["best_bytes"] == want["exhaustive"]["best_bytes"]
    # no real file is like the generated blocks: the gate sends every one of them to exhaustive best-per-block,
    # through the decision-site call too
    assert rep["router"]["fallback_blocks"] == len(blocks) and rep["router"]["loss_pct"] == 0.0
    dec, _ = R.route([b[1] for b in blocks], P, th, protos, radius)
    assert np.all(dec == -1)


@pytest.mark.skipif(not (os.path.exists(LABELS) and os.path.exists(FEATURES) and os.path.exists(R.MODEL)),
                    reason="router labels, features or model not present")
def test_gpu_router_route_at_the_decision_site(gpu):
    """kbwt_router_route, the one call kolmo's level-8 decision would make for a frame's host blocks, gives the
    decisions of features + network + threshold, on held-out blocks of every kind (and an empty and a one-byte
    block)."""
    lab = dict(np.load(LABELS))
    F = np.load(FEATURES)["features"].astype(np.float32)
    P, th = R.load_model()
    protos, radius = R.load_gate()
    idx = np.flatnonzero(np.arange(len(F)) % 5 == 4)[:40]
    blocks = [R.workload_block(lab["kind"][i], lab["seed"][i], lab["n"][i], model_seed=5) for i in idx]
    dec, probs = R.route(blocks + [np.zeros(0, np.uint8), np.array([7], np.uint8)], P, th, protos, radius)
    want = R.decide(R.predict(F[idx], P), th, R.gate(F[idx], protos), radius)
    say("route:", dec[:16], "want", want[:16])
    assert np.array_equal(dec[:len(idx)], want)
    assert np.allclose(probs[:len(idx)], R.predict(F[idx], P), atol=1e-6)
    assert probs.shape == (len(idx) + 2, R.NC) and np.allclose(probs.sum(1), 1, atol=1e-5)
