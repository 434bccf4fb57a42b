"""Shapes of the first sort's key space (GPU): runs of one byte inside random bytes (a run of R bytes gives about R rows
whose first keys agree, one group the doubling rounds split), whole blocks of one or two byte values, keys of five
8-bit digits (20- and 30-symbol alphabets, 5-bit codes), and such blocks mixed in one batch, where every block keeps its
own code width; key widths of five, six and seven 8-bit digits; blocks of every code width from 1 to 8 bits at lengths
around the m-gram histograms' chunk (16 Ki positions) and unit (128 Ki), and tiny ones whose end corrections cover the
whole block. Bar: bit-exact against the oracle, and every block round-trips.
"""
import sys

import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu
CAP = 2048   # run lengths around a workgroup's tile of 2048 rows


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


def alphabet_of(n, sigma, seed):
    """n bytes over sigma byte values spread over 0..255 (code width ceil(log2(sigma)) when all occur)."""
    rng = np.random.default_rng(seed)
    vals = np.sort(rng.permutation(256)[:sigma]).astype(np.uint8)
    return vals[rng.integers(0, sigma, n)]


def cases():
    yield "one byte", np.full(100000, 7, np.uint8)
    yield "two bytes", (np.random.default_rng(3).integers(0, 2, 300000) + 48).astype(np.uint8)
    for run in (CAP + 2, CAP + 3, CAP + 4, CAP + 5, 2 * CAP + 10, 5 * CAP):
        yield f"run {run}", with_run(300000, run, run)
    yield "runs at tile edges", np.concatenate([with_run(2 * CAP + 7, CAP + 3, 11, at=5), with_run(CAP * 3, CAP + 4, 12,
                                                                                                   at=CAP - 9)])
    yield "20 symbols", small_alphabet(600000, 20, 4)
    yield "64 symbols", small_alphabet(500000, 64, 6)
    yield "30 symbols", small_alphabet(300000, 30, 5)
    for kind, n in [("repeats", 1 << 20), ("dna", 700000), ("mixed", 700000), ("logs", 300000), ("csvcols", 300000),
                    ("random", 200000)]:
        yield kind, K.generate(kind, 61, n)


def test_gpu_first_sort_shapes(gpu):
    for tag, T in cases():
        say("first sort", tag, T.size)
        check(T, tag)


def test_gpu_first_sort_batch(gpu):
    """Blocks of different key widths and run shapes in one batch (groups never cross a block)."""
    blocks = [small_alphabet(200000, 20, 8), with_run(150000, CAP + 4, 9), np.full(5000, 3, np.uint8),
              K.generate("mixed", 62, 300000), with_run(4100, CAP + 3, 10, at=0), K.generate("dna", 63, 90000),
              np.frombuffer(b"ab", np.uint8).copy(), small_alphabet(70000, 30, 12)]
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]
    say("first sort batch", [b.size for b in blocks])
    Us, Is = K.bwt_aux_batch(blocks, rs)
    for b, (T, r, U, I) in enumerate(zip(blocks, rs, Us, Is)):
        eU, eI = refs.oracle_bwt(T, r)
        assert np.array_equal(U, eU) and np.array_equal(I, eI), f"block {b}"
        assert np.array_equal(K.unbwt_aux(U, r, I), T), f"block {b} inverse"


GRAM_SIGMAS = (1, 2, 3, 4, 7, 16, 40, 64, 100, 128, 200, 256)   # code widths 1, 1, 2, 2, 3, 4, 6, 6, 7, 7, 8, 8


def test_gpu_first_sort_gram_widths(gpu):
    """The digit histograms from code m-grams, for every code width: chunk and unit edges, the end corrections."""
    for sigma in GRAM_SIGMAS:
        for n in (2, 3, 17, 65, 16383, 16385, 131071, 131073):
            T = alphabet_of(n, sigma, 1000 * sigma + n)
            say("gram widths", sigma, n)
            check(T, f"sigma {sigma} n {n}")
    for sigma in (4, 40, 100):
        T = alphabet_of(2 * 131072 + 5, sigma, sigma)
        say("gram widths", sigma, T.size)
        check(T, f"sigma {sigma} n {T.size}")


def test_gpu_first_sort_gram_batch(gpu):
    """Blocks of every code width packed at odd offsets in one batch (m-gram units that start unaligned), with 5-bit
    blocks (per-digit histograms) between them."""
    rng = np.random.default_rng(77)
    blocks = []
    for k, sigma in enumerate(GRAM_SIGMAS + (20, 30)):
        n = int(rng.integers(1000, 40000)) * 2 + 1 if k % 3 else 131072 + 2 * k + 1
        blocks.append(alphabet_of(n, sigma, 500 + k))
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]
    say("gram batch", [b.size for b in blocks])
    Us, Is = K.bwt_aux_batch(blocks, rs)
    for b, (T, r, U, I) in enumerate(zip(blocks, rs, Us, Is)):
        eU, eI = refs.oracle_bwt(T, r)
        assert np.array_equal(U, eU) and np.array_equal(I, eI), f"block {b}"
        assert np.array_equal(K.unbwt_aux(U, r, I), T), f"block {b} inverse"
