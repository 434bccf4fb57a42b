"""The router's CPU oracle (oracle/router_oracle.py) pinned by what the mathematics fixes, and the router's host logic
(labels, weights, decisions, accounting). CPU only.

- network gradient: central finite differences in float64, and the closed form p - onehot at the output bias;
- order-1 entropy: closed forms (a constant block, a period-2 block, fair coin flips, uniform bytes);
- doubling counts: the final order equals the C oracle's suffix array, and on blocks whose last byte is the unique
  smallest byte m_r = #{p : max(LCP[p], LCP[p+1]) >= h0 * 2^(r-1)} (LCP by Kasai's algorithm).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import router_oracle as RO  # noqa: E402
import refs  # noqa: E402
from kolmo_amd import router as R  # noqa: E402


def small_problem(seed, ns=7):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 0.3, RO.NPARAM)
    X = rng.random((ns, RO.NF))
    y = rng.integers(0, RO.NC, ns)
    w = rng.random(ns) + 0.5
    return P, X, y, w


def test_gradient_matches_finite_differences():
    P, X, y, w = small_problem(1)
    L, g = RO.loss_grad(P, X, y, w)
    assert abs(L - RO.loss(P, X, y, w)) < 1e-12
    rng = np.random.default_rng(2)
    # every output-layer parameter and a sample of the others
    o3 = RO.NPARAM - (RO.NH * RO.NC + RO.NC)
    coords = list(range(o3, RO.NPARAM)) + list(rng.choice(o3, 300, replace=False))
    eps = 1e-6
    for k in coords:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += eps
        Pm[k] -= eps
        fd = (RO.loss(Pp, X, y, w) - RO.loss(Pm, X, y, w)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_output_bias_gradient_closed_form():
    P, X, y, w = small_problem(3, ns=11)
    p, _ = RO.forward(P, X)
    onehot = np.eye(RO.NC)[y]
    expect = np.mean(w[:, None] * (p - onehot), axis=0)
    _, g = RO.loss_grad(P, X, y, w)
    assert np.allclose(g[-RO.NC:], expect, atol=1e-14)


def test_order1_entropy_closed_forms():
    assert RO.order1_entropy(np.full(1000, 7, np.uint8)) == 0.0
    assert RO.order1_entropy(np.tile(np.array([1, 2], np.uint8), 500)) == 0.0
    rng = np.random.default_rng(4)
    coin = rng.integers(0, 2, 200000).astype(np.uint8)
    assert abs(RO.order1_entropy(coin) - 1.0) < 1e-3
    uni = rng.integers(0, 256, 4 << 20).astype(np.uint8)
    # plug-in estimate over 65536 bigram cells: biased low by about (cells - rows) / (2 ln 2 N)
    h = RO.order1_entropy(uni)
    assert 7.98 < h <= 8.0


def kasai(T, SA):
    n = len(T)
    rank = np.empty(n, np.int64)
    rank[SA] = np.arange(n)
    lcp = np.zeros(n + 1, np.int64)   # lcp[p]: between SA[p-1] and SA[p]; lcp[0] = lcp[n] = 0
    h = 0
    Tl = T.tolist()
    SAl = SA.tolist()
    for i in range(n):
        r = rank[i]
        if r > 0:
            j = SAl[r - 1]
            while i + h < n and j + h < n and Tl[i + h] == Tl[j + h]:
                h += 1
            lcp[r] = h
            if h:
                h -= 1
        else:
            h = 0
    return lcp


def sentinel_blocks():
    import kolmo_amd as K
    out = []
    for kind, n in [("markov2", 30000), ("repeats", 40000), ("dna", 30000), ("logs", 20000), ("csvcols", 20000)]:
        T = K.generate(kind, 11, n)
        out.append((kind, np.concatenate([np.maximum(T, 1), np.zeros(1, np.uint8)])))   # unique smallest last byte
    out.append(("abab", np.concatenate([np.tile(np.array([2, 3], np.uint8), 5000), np.ones(1, np.uint8)])))
    return out


@pytest.mark.parametrize("idx", range(6))
def test_doubling_counts_against_lcp(idx):
    kind, T = sentinel_blocks()[idx]
    act, order = RO.doubling(T, rounds=5)
    SA = refs.oracle_sa(T)
    assert np.array_equal(order, SA), kind
    lcp = kasai(T, SA.astype(np.int64))
    _, _, h0 = RO.codes_of(T)
    near = np.maximum(lcp[:-1], lcp[1:])   # row p against both neighbours
    for r in range(1, 6):
        assert act[r - 1] == int(np.sum(near >= h0 << (r - 1))), (kind, r)


def test_doubling_order_on_plain_blocks():
    import kolmo_amd as K
    for kind, n in [("mixed", 50000), ("random", 20000), ("repeats", 60000), ("bf16", 30000)]:
        T = K.generate(kind, 12, n)
        _, order = RO.doubling(T)
        assert np.array_equal(order, refs.oracle_sa(T)), kind


def test_labels_weights_decisions():
    n = np.array([1000, 1000, 1000, 1000, 1000])
    sizes = R.outcome_sizes(n, [500, 700, -1, 300, 400], [600, 650, -1, 300, 500], [520, 720, -1, 300, 380])
    assert sizes.tolist() == [[500, 600, 1000, 520], [700, 650, 1000, 720], [1000, 1000, 1000, 1000],
                              [300, 300, 1000, 300], [400, 500, 1000, 380]]
    # ties go to storing, then the whole block's sorting, then KLZMA, then the halves
    assert R.labels(sizes).tolist() == [0, 1, 2, 0, 3]
    w = R.weights(sizes)
    assert abs(w.mean() - 1) < 1e-6 and w[0] > w[1]   # 0: choosing wrong costs 20-500 bytes, 1: 50-350
    probs = np.array([[0.9, 0.05, 0.03, 0.02], [0.4, 0.35, 0.2, 0.05], [0.2, 0.1, 0.6, 0.1], [0.1, 0.8, 0.05, 0.05],
                      [0.1, 0.1, 0.1, 0.7]])
    d = R.decide(probs, 0.5)
    assert d.tolist() == [0, -1, 2, 1, 3]
    one, four, two = np.ones(5), np.full(5, 4.0), np.full(5, 2.0)
    ev = R.evaluate(d, sizes, one, four, two)
    assert ev["bytes"] == 500 + 650 + 1000 + 300 + 380 and ev["best_bytes"] == ev["bytes"]
    assert ev["seconds"] == 1 + (1 + 4 + 2) + 0 + 4 + 2 and ev["fallback_blocks"] == 1
    # the gate: a block farther than the radius from every prototype is coded every way, whatever its probability
    g = R.decide(probs, 0.5, dist=np.array([0.1, 0.1, 0.3, 0.2, 0.25]), radius=0.2)
    assert g.tolist() == [0, -1, -1, 1, -1]


def test_prototypes_farthest_point_sampling():
    """Farthest-point sampling on points of a line: from 0 it picks the far end, then the middle, then the quarter
    points; the coverage radius is the largest distance from a point to its nearest pick."""
    X = np.zeros((9, 3))
    X[:, 0] = np.arange(9)
    pick, cover = R.prototypes(X, 2)
    assert pick.tolist() == [0, 8] and cover == 4.0
    pick, cover = R.prototypes(X, 3)
    assert pick.tolist() == [0, 8, 4] and cover == 2.0
    pick, cover = R.prototypes(X, 5)
    assert sorted(pick.tolist()) == [0, 2, 4, 6, 8] and cover == 1.0
    pick, cover = R.prototypes(X, 50)
    assert sorted(pick.tolist()) == list(range(9)) and cover == 0.0
    # two clusters: two prototypes land one in each, and the coverage is a cluster's own spread
    rng = np.random.default_rng(3)
    Y = np.concatenate([rng.normal(0, 0.01, (50, 4)), rng.normal(5, 0.01, (50, 4))])
    pick, cover = R.prototypes(Y, 2)
    assert {int(p) // 50 for p in pick} == {0, 1} and cover < 0.1
