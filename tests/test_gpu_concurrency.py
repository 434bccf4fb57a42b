"""Concurrent calls of kolmo_sais_bwt_aux / kolmo_sais_unbwt_aux from several host threads, as kolmo's block workers
make them (src/core/compress.c:231-256, src/core/decode.c:404-410): the library gathers them into batches, and every
block still gets the oracle's result. Forward and inverse calls run at the same time, so batches of both kinds form
while the other kind is queued."""
import sys
import threading
import time

import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def blocks_for(seed, count):
    kinds = ["markov2", "mixed", "dna", "logs", "random", "repeats", "csvcols", "bf16"]
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(2, 900000, count)]
    sizes[0], sizes[1] = 2, 65537                 # the smallest block and one over 64 KiB
    return [K.generate(kinds[i % len(kinds)], seed + i, n) for i, n in enumerate(sizes)]


def run_threads(fns):
    errors = []

    def wrap(f):
        try:
            f()
        except BaseException as e:   # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if errors:
        raise errors[0]
    return time.perf_counter() - t0


def test_gpu_concurrent_calls_equal_oracle(gpu):
    blocks = blocks_for(900, 24)
    rs = [1 << K.codec18_rlog(T.size) for T in blocks]
    expect = [refs.oracle_bwt(T, r) for T, r in zip(blocks, rs)]
    say("concurrency: oracle done for", len(blocks), "blocks,", sum(T.size for T in blocks), "bytes")
    got = [None] * len(blocks)
    back = [None] * len(blocks)

    def fwd(k):
        def f():
            for _ in range(3):
                got[k] = K.bwt_aux(blocks[k], rs[k])
        return f

    def inv(k):
        def f():
            eU, eI = expect[k]
            for _ in range(3):
                back[k] = K.unbwt_aux(eU, rs[k], eI)
        return f

    # 12 threads: each runs forwards of its blocks while others run inverses
    fns = [fwd(k) for k in range(0, len(blocks), 2)] + [inv(k) for k in range(1, len(blocks), 2)]
    secs = run_threads(fns)
    say("concurrency: 12 threads, 3 calls each, in", round(secs, 3), "s")
    for k in range(0, len(blocks), 2):
        assert np.array_equal(got[k][0], expect[k][0]) and np.array_equal(got[k][1], expect[k][1]), k
    for k in range(1, len(blocks), 2):
        assert np.array_equal(back[k], blocks[k]), k
    # and the other way round, with r == n (kolmo_sais_unbwt: one index) for the inverses
    got = [None] * len(blocks)
    back = [None] * len(blocks)

    def fwd2(k):
        def f():
            got[k] = K.bwt_aux(blocks[k], rs[k])
        return f

    def inv_one(k):
        def f():
            eU, eI = expect[k]
            back[k] = K.unbwt_aux(eU, blocks[k].size, eI[:1])   # r == n: the primary index only
        return f

    run_threads([fwd2(k) for k in range(1, len(blocks), 2)] + [inv_one(k) for k in range(0, len(blocks), 2)])
    for k in range(1, len(blocks), 2):
        assert np.array_equal(got[k][0], expect[k][0]) and np.array_equal(got[k][1], expect[k][1]), k
    for k in range(0, len(blocks), 2):
        assert np.array_equal(back[k], blocks[k]), k


def test_gpu_concurrent_errors_stay_per_call(gpu):
    """An invalid call among concurrent ones fails alone; the valid ones are unaffected."""
    blocks = blocks_for(950, 8)
    rs = [1 << K.codec18_rlog(T.size) for T in blocks]
    expect = [refs.oracle_bwt(T, r) for T, r in zip(blocks, rs)]
    out, errs = [None] * len(blocks), [None] * len(blocks)

    def call(k):
        def f():
            try:
                if k == 3:
                    eU, eI = expect[k]
                    bad = eI.copy()
                    bad[0] = blocks[k].size + 1          # a primary index out of range: -1 before any queueing
                    out[k] = K.unbwt_aux(eU, rs[k], bad)
                else:
                    out[k] = K.bwt_aux(blocks[k], rs[k])
            except K.KbwtError as e:
                errs[k] = e.code
        return f

    run_threads([call(k) for k in range(len(blocks))])
    assert errs[3] == -1 and out[3] is None
    for k in range(len(blocks)):
        if k != 3:
            assert errs[k] is None
            assert np.array_equal(out[k][0], expect[k][0]) and np.array_equal(out[k][1], expect[k][1]), k


def test_gpu_concurrent_large_calls(gpu):
    """Concurrent calls whose batch holds more than a few MiB of host bytes: the blocks cross the boundaries of the
    library's pinned staging chunks, in both directions, and every block still gets the oracle's result."""
    kinds = ["markov2", "mixed", "dna", "logs", "random", "repeats"]
    rng = np.random.default_rng(960)
    sizes = [int(x) for x in rng.integers(1 << 20, 5 << 19, 12)]
    sizes[0] = (1 << 21) + 1
    blocks = [K.generate(kinds[i % len(kinds)], 960 + i, n) for i, n in enumerate(sizes)]
    rs = [1 << K.codec18_rlog(T.size) for T in blocks]
    expect = [refs.oracle_bwt(T, r) for T, r in zip(blocks, rs)]
    say("concurrency: oracle done for", len(blocks), "blocks,", sum(sizes), "bytes")
    got = [None] * len(blocks)
    back = [None] * len(blocks)

    def fwd(k):
        def f():
            got[k] = K.bwt_aux(blocks[k], rs[k])
        return f

    def inv(k):
        def f():
            eU, eI = expect[k]
            back[k] = K.unbwt_aux(eU, rs[k], eI)
        return f

    run_threads([fwd(k) for k in range(len(blocks))])
    run_threads([inv(k) for k in range(len(blocks))])
    for k in range(len(blocks)):
        assert np.array_equal(got[k][0], expect[k][0]) and np.array_equal(got[k][1], expect[k][1]), k
        assert np.array_equal(back[k], blocks[k]), k


def test_gpu_block_workers_from_c_threads(gpu):
    """kolmo's block workers as C threads (no interpreter lock between calls): 12 threads over 6.5 MiB in 1 MiB
    blocks and a ragged last one, forward then inverse per block, every round trip equal to its block. One thread
    over the same blocks as well."""
    T = K.generate("mixed", 970, (13 << 19) + 333)
    for workers in (12, 1):
        sec = K.block_workers(T, 1 << 20, workers, reps=1)
        say("block workers:", workers, "threads,", round(T.size / sec / 2**20, 1), "MiB/s")
        assert sec > 0
