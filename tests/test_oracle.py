"""The oracle pinned before it is trusted (CPU only).

- oracle_sa against brute-force suffix sorting (Python's bytes order is exactly the implicit-end-marker order: a
  proper prefix sorts first), exhaustively over {a,b}^<=12 and on random strings;
- oracle_bwt_aux / oracle_unbwt_aux against kolmo's own sais.c compiled from source (oracle/_ref), on the edge plan of
  kolmo's tests/test_gpu_sa.c:74-98 and random inputs with long repeats;
- the codec-5 stream over the oracle against the 63 libbsc 3.4.0 golden hashes of tests/test_bwt_golden.h;
- the committed fixtures (written from oracle/_ref) and the sampling-interval rules against their closed forms.
A planted mistake (a dropped sentinel rule, a swapped primary row, an off-by-one sample) fails at least one of these.
"""
import itertools
import json
import os

import numpy as np
import pytest

import kolmo_amd as K
import refs

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not refs.have_ref(), reason="oracle/_ref not built (needs the kolmo tree once)")


def brute_sa(T):
    b = bytes(T)
    return sorted(range(len(b)), key=lambda i: b[i:])


def brute_bwt(T, r):
    """libsais' BWT convention from the definition (sais.c:12043-12049, metal.h:23-25)."""
    n = len(T)
    sa = brute_sa(T)
    p0 = sa.index(0)
    U = [T[n - 1]] + [T[sa[p] - 1] for p in range(n) if p != p0]
    I = [sa.index(k * r) + 1 for k in range((n - 1) // r + 1)]
    return bytes(U), I


def test_oracle_sa_exhaustive_binary():
    for n in range(1, 13):
        for bits in itertools.product(b"ab", repeat=n):
            T = bytes(bits)
            assert list(refs.oracle_sa(T)) == brute_sa(T), T


def test_oracle_sa_random():
    rng = np.random.default_rng(1)
    for it in range(200):
        n = int(rng.integers(1, 2048))
        alpha = int(rng.choice([1, 2, 3, 4, 17, 256]))
        T = rng.integers(0, alpha, n, dtype=np.uint8)
        if it % 3 == 0 and n > 100:   # periodic stretches
            p = int(rng.integers(1, 30))
            T[50:] = np.resize(T[:p], n - 50)
        assert list(refs.oracle_sa(T)) == brute_sa(T)


def test_oracle_bwt_bruteforce_small():
    rng = np.random.default_rng(2)
    for it in range(300):
        n = int(rng.integers(2, 200))
        T = bytes(rng.integers(0, int(rng.choice([1, 2, 3, 256])), n, dtype=np.uint8))
        r = 1 << int(rng.integers(1, 8))
        U, I = refs.oracle_bwt(T, r)
        eu, ei = brute_bwt(T, r)
        assert bytes(U) == eu and list(I) == ei
        assert bytes(refs.oracle_unbwt(U, r, I)) == T
        assert bytes(refs.oracle_unbwt(U, n, I[:1])) == T


def test_oracle_validation():
    O = refs.oracle()
    T = np.zeros(10, np.uint8)
    U = np.zeros(10, np.uint8)
    I = np.zeros(10, np.int32)
    p = refs._p
    assert O.oracle_bwt_aux(p(T), p(U), 10, 3, p(I, refs._i32p)) == -1   # r not a power of two
    assert O.oracle_bwt_aux(p(T), p(U), 10, 1, p(I, refs._i32p)) == -1
    assert O.oracle_bwt_aux(p(T), p(U), -1, 2, p(I, refs._i32p)) == -1
    I[:] = 11
    assert O.oracle_unbwt_aux(p(T), p(U), 10, 2, p(I, refs._i32p)) == -1  # index out of range


def edge_inputs():
    """tests/test_gpu_sa.c:74-98: n = 2..299 x 7 kinds (zeros, one symbol, period 2, "abcab", random binary, half zeros,
    random bytes), with the sampling interval of that test (largest power of two <= n/8, at least 2)."""
    rng = np.random.default_rng(3)
    for n in range(2, 300):
        for kind in range(7):
            i = np.arange(n)
            if kind == 0: T = np.zeros(n, np.uint8)
            elif kind == 1: T = np.full(n, ord("a"), np.uint8)
            elif kind == 2: T = (i % 2).astype(np.uint8)
            elif kind == 3: T = np.frombuffer(b"abcab", np.uint8)[i % 5]
            elif kind == 4: T = rng.integers(0, 2, n, dtype=np.uint8)
            elif kind == 5: T = np.where(i < n // 2, 0, rng.integers(0, 3, n)).astype(np.uint8)
            else: T = rng.integers(0, 256, n, dtype=np.uint8)
            r = (n // 8) | 1
            while r & (r - 1):
                r &= r - 1
            yield kind, n, np.ascontiguousarray(T), max(r, 2)


def random_inputs(count=30, big=300000, seed=4):
    """tests/test_gpu_sa.c:89-98: random alphabets, and long repeats (copies of an earlier stretch)."""
    rng = np.random.default_rng(seed)
    for it in range(count):
        n = 2 + int(rng.integers(0, 70000 if it < 2 * count // 3 else big))
        alpha = 1 + int(rng.integers(0, 3 if it % 4 == 0 else 256))
        T = rng.integers(0, alpha, n, dtype=np.uint8)
        if it % 5 == 1 and n > 1000:
            keep = rng.integers(0, 1000, n - 1000) != 0
            for i in range(1000, n):   # copy of T[i - 997] except 1 in 1000
                if keep[i - 1000]:
                    T[i] = T[i - 997]
        yield it, T


@needs_ref
def test_oracle_matches_reference_edge_plan():
    R = refs.ref()
    for kind, n, T, r in edge_inputs():
        U, I = refs.oracle_bwt(T, r)
        rc, U2, I2 = refs.sais_bwt(R, T, r)
        assert rc == 0 and np.array_equal(U, U2) and np.array_equal(I, I2), (kind, n)
        rc, T2 = refs.sais_unbwt(R, U, r, I)
        assert rc == 0 and np.array_equal(T2, T)
        assert np.array_equal(refs.oracle_unbwt(U, r, I), T)


@needs_ref
def test_oracle_matches_reference_random():
    R = refs.ref()
    for it, T in random_inputs():
        n = T.size
        r = 1 << K.codec18_rlog(n)
        U, I = refs.oracle_bwt(T, r)
        rc, U2, I2 = refs.sais_bwt(R, T, r)
        assert rc == 0 and np.array_equal(U, U2) and np.array_equal(I, I2), it
        assert np.array_equal(refs.oracle_unbwt(U, r, I), T)


@needs_ref
def test_oracle_codec5_golden_hashes():
    """The oracle under kolmo's codec-5 glue writes the bytes libbsc 3.4.0 wrote (tests/test_bwt_golden.h)."""
    g = json.load(open(os.path.join(HERE, "golden", "bsc_golden.json")))
    C = refs.codec("oracle")
    for kind in range(9):
        for j, n in enumerate(g["sizes"]):
            T = K.generate_bsc_kind(kind, n)
            out = refs.bsc_compress(C, T, g["lzp_hash"], g["lzp_min"], g["coder"])
            assert f"{refs.fnv1a(out):016x}" == g["fnv1a"][kind][j], (kind, n)


def test_golden_inputs_reproduce():
    """The generators still produce the inputs the fixtures were written from."""
    fx = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["inputs"]
    for e in fx:
        if e["n"] > (8 << 20):
            continue
        assert refs.sha256(K.generate(e["gen"], e["seed"], e["n"])) == e["sha_T"], e["name"]


def test_oracle_matches_fixtures():
    fx = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["inputs"]
    for e in fx:
        if e["n"] > (8 << 20):
            continue
        T = K.generate(e["gen"], e["seed"], e["n"])
        U, I = refs.oracle_bwt(T, e["r"])
        assert refs.sha256(U) == e["sha_U"] and refs.sha256(I.astype("<i4")) == e["sha_I"], e["name"]
        assert int(I[0]) == e["primary"]


def test_sampling_rules_closed_form():
    O = refs.oracle()
    for n in list(range(2, 5000)) + [65535, 65536, 1 << 20, (1 << 20) + 7, 32 << 20, 128 << 20, 0x7FFFFFF0]:
        x = n // 8
        expect_mod = (1 << (x.bit_length() - 1)) - 1 if x > 0 else 0
        assert K.bwt_mod(n) == expect_mod == O.oracle_bwt_mod(n)
        rl = max(1, (expect_mod + 1).bit_length() - 1)
        while rl < 30 and (n - 1) >> rl > 255:
            rl += 1
        assert K.codec18_rlog(n) == rl == O.oracle_codec18_rlog(n)
        assert (n - 1) >> rl <= 255


@needs_ref
def test_sampling_rule_matches_codec18_header():
    """The rlog and k that kolmo's own codec-18 encoder writes (bwt_stream.c:224-262) follow codec18_rlog."""
    R = refs.ref()
    for n in [1000, 4096, 65536, 300000, 1 << 20]:
        T = K.generate_bsc_kind(8, n)   # word text: always compressible
        rc, S = refs.bwt18_compress(R, T, params=(0, 0, 2, 0, 0, 0))   # no LZP: the sort sees all n bytes
        assert rc > 0
        flags = int(S[0])
        assert not flags & 4
        rlog, k = int(S[5]), int(S[6])
        assert rlog == K.codec18_rlog(n) and k == (n - 1) >> rlog


def test_batch_cfg4_fixture_pinned():
    """tests/golden/batch_cfg4.json (the benchmark's 512-block batch, written from kolmo's own C): the generator
    reproduces every block, and the oracle gives the recorded BWT and indexes for a sample of them."""
    import hashlib
    import json
    fx = json.load(open(os.path.join(HERE, "golden", "batch_cfg4.json")))
    blocks, _, _ = K.generate_batch(fx["seed"], fx["count"], fx["lo"], fx["hi"], fx["kinds"])
    assert len(blocks) == fx["count"] == len(fx["blocks"])
    for T, e in zip(blocks, fx["blocks"]):
        assert T.size == e["n"] and refs.sha256(T) == e["sha_T"]
    small = sorted(range(len(blocks)), key=lambda b: blocks[b].size)
    for b in small[:6] + small[200:203] + [small[-1]]:
        e = fx["blocks"][b]
        U, I = refs.oracle_bwt(blocks[b], e["r"])
        h = hashlib.sha256(U.tobytes())
        h.update(np.asarray(I, "<i4").tobytes())
        assert h.hexdigest() == e["sha_UI"], (b, e["kind"], e["n"])
