"""The end of the first sort (GPU): the rows that share the top digits of their first keys form buckets, which are
sorted in LDS by the low digits (the default), with buckets of more than 2048 rows left to the global radix path; or
every digit goes through the global radix passes (KBWT_FIRST_LDS=0). Both must give the reference's bytes.

Inputs put buckets at the LDS limit and just past it (runs of one byte inside random bytes: a run of R bytes gives a
bucket of R - 3 rows whose first four bytes agree), make a whole block one bucket, use keys of five digits (20-symbol
alphabets, low digits 2 instead of 3), and mix such blocks in one batch. Bar: bit-exact against the oracle, and every
block round-trips.
"""
import sys

import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu
CAP = 2048   # rows of the largest bucket sorted in LDS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(T, tag):
    r = 1 << K.codec18_rlog(T.size)
    U, I = K.bwt_aux(T, r)
    eU, eI = refs.oracle_bwt(T, r)
    assert np.array_equal(U, eU), f"{tag}: BWT bytes differ"
    assert np.array_equal(I, eI), f"{tag}: sampled indexes differ"
    assert np.array_equal(K.unbwt_aux(U, r, I), T), f"{tag}: inverse"


def with_run(n, run, seed, at=None, byte=97):
    rng = np.random.default_rng(seed)
    T = rng.integers(0, 256, n, dtype=np.uint8)
    at = n // 3 if at is None else at
    T[at:at + run] = byte
    return T


def small_alphabet(n, sigma, seed):
    rng = np.random.default_rng(seed)
    p = rng.dirichlet(np.full(sigma, 0.5))
    return (rng.choice(sigma, n, p=p) + 65).astype(np.uint8)


def cases():
    yield "one byte", np.full(100000, 7, np.uint8)
    yield "two bytes", (np.random.default_rng(3).integers(0, 2, 300000) + 48).astype(np.uint8)
    for run in (CAP + 2, CAP + 3, CAP + 4, CAP + 5, 2 * CAP + 10, 5 * CAP):
        yield f"run {run}", with_run(300000, run, run)
    yield "runs at tile edges", np.concatenate([with_run(2 * CAP + 7, CAP + 3, 11, at=5), with_run(CAP * 3, CAP + 4, 12,
                                                                                                   at=CAP - 9)])
    yield "20 symbols", small_alphabet(600000, 20, 4)
    yield "30 symbols", small_alphabet(300000, 30, 5)
    for kind, n in [("repeats", 1 << 20), ("dna", 700000), ("mixed", 700000), ("logs", 300000), ("csvcols", 300000),
                    ("random", 200000)]:
        yield kind, K.generate(kind, 61, n)


@pytest.mark.parametrize("lds", ["1", "0"])
def test_gpu_first_sort_paths(gpu, lds, monkeypatch):
    monkeypatch.setenv("KBWT_FIRST_LDS", lds)
    for tag, T in cases():
        say("first sort", lds, tag, T.size)
        check(T, f"{tag} (KBWT_FIRST_LDS={lds})")


@pytest.mark.parametrize("lds", ["1", "0"])
def test_gpu_first_sort_batch(gpu, lds, monkeypatch):
    """Blocks of different key widths and bucket shapes in one batch (a bucket never crosses a block)."""
    monkeypatch.setenv("KBWT_FIRST_LDS", lds)
    blocks = [small_alphabet(200000, 20, 8), with_run(150000, CAP + 4, 9), np.full(5000, 3, np.uint8),
              K.generate("mixed", 62, 300000), with_run(4100, CAP + 3, 10, at=0), K.generate("dna", 63, 90000),
              np.frombuffer(b"ab", np.uint8).copy(), small_alphabet(70000, 30, 12)]
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]
    say("first sort batch", lds, [b.size for b in blocks])
    Us, Is = K.bwt_aux_batch(blocks, rs)
    for b, (T, r, U, I) in enumerate(zip(blocks, rs, Us, Is)):
        eU, eI = refs.oracle_bwt(T, r)
        assert np.array_equal(U, eU) and np.array_equal(I, eI), f"block {b} (KBWT_FIRST_LDS={lds})"
        assert np.array_equal(K.unbwt_aux(U, r, I), T), f"block {b} inverse"
