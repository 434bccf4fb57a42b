"""CPU oracle of the learned router (test infrastructure: the checker, never the product).

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg may import this module. The router has no counterpart in
kolmo; what it restates is textbook mathematics, in float64 numpy:

  - the network: 262 -> 64 -> 64 -> 4 with ReLU and softmax, parameters in the library's input-major layout
    (W1[k*64+j], b1, W2[k*64+j], b2, W3[k*4+c], b3), and the gradient of the weighted mean cross-entropy by
    back-propagation. Pinned by central finite differences (tests/test_router_oracle.py) and by the closed form of the
    softmax cross-entropy gradient (p - onehot) at the logits;
  - the features: sqrt(byte count / n), the order-1 entropy H(T[i] | T[i-1]) / 8 from bigram counts (pinned by
    closed forms: a constant block, a period-2 block, i.i.d. bytes), and m_r / n, the share of suffixes still in a
    rank group at the start of doubling round r = 1..5. The doubling restated here follows the library's conventions
    (dense symbol codes of w = ceil(log2 sigma) bits, h0 symbols in the first key, zero past the end; round r offset
    h0 * 2^(r-1), capped at n; past the end a suffix's second key is n - 1 - i, below every rank). It is pinned by
    its final order, which must be the suffix array of the C oracle (oracle/kbwt_oracle.c), and by the LCP form of
    the counts: on a block whose last byte is its unique smallest byte (so zero padding never matches a real
    symbol), m_r = #{p : max(LCP[p], LCP[p+1]) >= h0 * 2^(r-1)} exactly, with LCP from Kasai's algorithm over the
    C oracle's suffix array.
"""
import numpy as np

NF, NH, NC = 262, 64, 4
NPARAM = NF * NH + NH + NH * NH + NH + NH * NC + NC
ROUNDS = 5


def unpack(P):
    P = np.asarray(P, np.float64)
    o = 0
    W1 = P[o: o + NF * NH].reshape(NF, NH); o += NF * NH
    b1 = P[o: o + NH]; o += NH
    W2 = P[o: o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = P[o: o + NH]; o += NH
    W3 = P[o: o + NH * NC].reshape(NH, NC); o += NH * NC
    b3 = P[o: o + NC]; o += NC
    assert o == NPARAM
    return W1, b1, W2, b2, W3, b3


def forward(P, X):
    """Softmax probabilities and the intermediate activations."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ W3 + b3
    zm = z - z.max(axis=1, keepdims=True)
    e = np.exp(zm)
    p = e / e.sum(axis=1, keepdims=True)
    return p, (X, a1, h1, a2, h2, z)


def loss(P, X, y, w=None):
    p, _ = forward(P, X)
    y = np.asarray(y)
    w = np.ones(len(y)) if w is None else np.asarray(w, np.float64)
    return float(np.mean(w * -np.log(p[np.arange(len(y)), y])))


def loss_grad(P, X, y, w=None):
    """Mean over samples of w * cross-entropy, and its gradient in the flat parameter layout."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    p, (X, a1, h1, a2, h2, z) = forward(P, X)
    ns = X.shape[0]
    y = np.asarray(y)
    w = np.ones(ns) if w is None else np.asarray(w, np.float64)
    L = float(np.mean(w * -np.log(p[np.arange(ns), y])))
    dz = p.copy()
    dz[np.arange(ns), y] -= 1.0
    dz *= (w / ns)[:, None]
    gW3 = h2.T @ dz
    gb3 = dz.sum(0)
    dh2 = (dz @ W3.T) * (a2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


# ---- features

def codes_of(T):
    """Dense order-preserving symbol codes, their width w and the first key's symbol count h0."""
    T = np.asarray(T, np.uint8)
    present = np.bincount(T, minlength=256) > 0
    sigma = int(present.sum())
    code = np.cumsum(present) - 1
    w = 1 if sigma <= 2 else int(np.ceil(np.log2(sigma)))
    h0 = min(8, 56 // w) if w >= 5 else 56 // w
    return code[T].astype(np.int64), w, h0


def doubling(T, rounds=ROUNDS):
    """Prefix doubling with the library's conventions. Returns (active, order): active[r-1] = suffixes in groups of
    two or more at the start of round r (r = 1..rounds), order = the final suffix order."""
    n = len(T)
    c, w, h0 = codes_of(T)
    key = np.zeros(n, np.int64)
    pad = np.concatenate([c, np.zeros(h0, np.int64)])
    for k in range(h0):
        key = (key << w) | pad[k: k + n]
    order = np.argsort(key, kind="stable")
    sk = key[order]
    start = np.r_[True, sk[1:] != sk[:-1]]
    head_of_row = np.maximum.accumulate(np.where(start, np.arange(n), 0))
    rank = np.empty(n, np.int64)
    rank[order] = head_of_row
    active = []
    r = 0
    while True:
        r += 1
        sizes = np.bincount(rank, minlength=n)
        in_group = sizes[rank] >= 2
        m = int(in_group.sum())
        if r <= rounds:
            active.append(m)
        if m == 0:
            break
        h = min(h0 << (r - 1), n)
        i = np.arange(n)
        j = i + h
        sec = np.where(j < n, rank[np.minimum(j, n - 1)] + h, n - 1 - i)
        order = np.lexsort((sec, rank))
        k1, k2 = rank[order], sec[order]
        start = np.r_[True, (k1[1:] != k1[:-1]) | (k2[1:] != k2[:-1])]
        head_of_row = np.maximum.accumulate(np.where(start, np.arange(n), 0))
        rank = np.empty(n, np.int64)
        rank[order] = head_of_row
    active += [0] * (rounds - len(active))
    return active, order


def order1_entropy(T):
    T = np.asarray(T, np.uint8)
    if T.size < 2:
        return 0.0
    pairs = T[:-1].astype(np.int64) * 256 + T[1:]
    cnt = np.bincount(pairs, minlength=65536).reshape(256, 256).astype(np.float64)
    cy = cnt.sum(axis=1, keepdims=True)
    nz = cnt > 0
    h = np.sum(cnt[nz] * (np.log2(np.broadcast_to(cy, cnt.shape)[nz]) - np.log2(cnt[nz])))
    return float(h / (T.size - 1))


def features(T):
    T = np.asarray(T, np.uint8)
    n = T.size
    f = np.zeros(NF, np.float64)
    f[:256] = np.sqrt(np.bincount(T, minlength=256) / n)
    f[256] = order1_entropy(T) / 8.0
    act, _ = doubling(T)
    f[257:] = np.asarray(act, np.float64) / n
    return f
