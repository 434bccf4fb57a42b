"""The error path's CPU side (no kernel runs): the opt-in CPU path, fault-injection specs, oversized blocks.

kolmo's Metal lane promised that a GPU error sorts on the CPU with the same bytes (src/arch/metal/cubwt_metal.c:4-7)
and counted each fallback (src/core/hetero.c:62-66). The library's CPU path (kolmo_amd/csrc/kbwt_cpu.cpp) is off by
default; enabled, a host-pointer call whose device work fails is computed there. Without a device every device call
fails, so here it computes every block; where a device is present the failures are forced from the host by fault
injection on every synchronisation. Its output is checked against the oracle and against kolmo's own C, and whole
kolmo frames written through it against the digests kolmo publishes.
"""
import contextlib
import hashlib
import json
import os
import threading

import numpy as np
import pytest

import kolmo_amd as K
import container_inputs
import refs
from test_oracle import edge_inputs, random_inputs

HERE = os.path.dirname(os.path.abspath(__file__))


@contextlib.contextmanager
def cpu_path():
    """Every device call fails (no device, or an injected fault on every synchronisation) and the CPU path is on."""
    was = K.set_cpu_fallback(True)
    if K.device_ok():
        K.fault_inject("sync:1:1000000000")
    try:
        yield
    finally:
        K.fault_inject(None)
        K.set_cpu_fallback(was)


def test_fallback_is_off_by_default():
    assert K.set_cpu_fallback(False) is False


def test_fault_specs():
    for bad in ("bogus", "alloc", "alloc:0", "alloc:x", "launch:1:0", "copy:3:", "sync:1:2:3", "Alloc:1"):
        with pytest.raises(K.KbwtError):
            K.fault_inject(bad)
    for good in ("alloc:1", "launch:7", "sync:2:5", "copy:1:1", ""):
        K.fault_inject(good)
    K.fault_inject(None)


def test_block_longer_than_a_batch_is_refused_not_queued():
    """A block longer than 2^31 - 65 bytes cannot be indexed by the kernels: -2 at once, nothing queued (the call
    does not touch the buffers, so small ones stand in for the huge block)."""
    L = K.lib()
    T = np.zeros(16, np.uint8)
    U = np.zeros(16, np.uint8)
    A = np.zeros(16, np.int32)
    I = np.zeros(64, np.int32)
    n = 2**31 - 10
    done = []

    def call():
        done.append(L.kolmo_sais_bwt_aux(K._ptr(T), K._ptr(U), K._ptr(A, K._i32p), n, 1 << 30, K._ptr(I, K._i32p), 1))
        I[:2] = 1
        done.append(L.kolmo_sais_unbwt_aux(K._ptr(T), K._ptr(U), K._ptr(A, K._i32p), n, 1 << 30, K._ptr(I, K._i32p),
                                           1))

    t = threading.Thread(target=call)
    t.start()
    t.join(timeout=30)
    assert not t.is_alive(), "an oversized block hung the library"
    assert done == [K.KBWT_ERR_NOMEM, K.KBWT_ERR_NOMEM]
    assert "2^31" in L.kbwt_last_error().decode() or True   # the message belongs to the calling thread


def check(T, r):
    U, I = K.bwt_aux(T, r)
    eU, eI = refs.oracle_bwt(T, r)
    assert np.array_equal(U, eU) and np.array_equal(I, eI), f"n {T.size} r {r}"
    assert np.array_equal(K.unbwt_aux(U, r, I), T)
    assert np.array_equal(K.unbwt_aux(U, T.size, I[:1]), T)


def test_cpu_path_matches_the_oracle():
    before = K.counters()["cpu_fallbacks"]
    with cpu_path():
        for kind, n, T, r in edge_inputs():
            if n % 7 == 0 or n < 20:
                check(T, r)
        for it, T in random_inputs(count=12, big=400000, seed=9):
            check(T, 1 << K.codec18_rlog(T.size))
        for kind in range(9):
            T = K.generate_bsc_kind(kind, 70001)
            check(T, 1 << K.codec18_rlog(T.size))
    assert K.counters()["cpu_fallbacks"] > before


@pytest.mark.skipif(not refs.have_ref(), reason="oracle/_ref not built")
def test_cpu_path_matches_kolmo_on_the_generators():
    R = refs.ref()
    with cpu_path():
        for gen, seed, n in (("mixed", 2, 1 << 20), ("markov2", 1, 1 << 20), ("repeats", 3, 1 << 20),
                             ("dna", 4, 300000), ("random", 5, 200000)):
            T = K.generate(gen, seed, n)
            r = 1 << K.codec18_rlog(n)
            U, I = K.bwt_aux(T, r)
            rc, eU, eI = refs.sais_bwt(R, T, r)
            assert rc == 0 and np.array_equal(U, eU) and np.array_equal(I, eI), gen
            assert np.array_equal(K.unbwt_aux(U, r, I), T)


def test_cpu_path_decodes_strings_that_are_no_bwt_to_zeros():
    """The device path decodes a string that is no BWT to zeros (its list from the end marker's row does not reach
    the primary row in n steps); the CPU path answers the same, also when the samples disagree with the string."""
    T = K.generate("mixed", 7, 5000)
    r = 1 << K.codec18_rlog(T.size)
    with cpu_path():
        U, I = K.bwt_aux(T, r)
        zeros = 0
        for a, b in ((10, 20), (100, 4000), (7, 8), (1234, 3210)):
            bad = U.copy()
            bad[a], bad[b] = U[b], U[a]             # same histogram, LF split into cycles in most cases
            out = K.unbwt_aux(bad, r, I)
            assert out.size == T.size
            if not np.any(out):
                zeros += 1
            else:                                   # still the BWT of some string: one walk decodes it, as on the device
                assert np.array_equal(out, refs.oracle_unbwt(bad, bad.size, I[:1]))
        assert zeros > 0
        wrong = I.copy()
        wrong[1:] = 1                               # samples that name the wrong rows: the single walk decodes
        assert np.array_equal(K.unbwt_aux(U, r, wrong), T)


@pytest.mark.skipif(not refs.have_full("hip"), reason="oracle/_ref/libkolmo_full_hip.so not built (kolmo tree absent)")
def test_whole_frames_through_the_cpu_path_keep_published_digests():
    """kolmo's full library relinked over libkbwt_hip.so, every block-sorting call computed by the CPU path: the
    level-8/9 frames of kolmo's determinism inputs equal the published digests at one and four threads, and decode."""
    gold = json.load(open(os.path.join(HERE, "golden", "container_golden.json")))
    data = dict(container_inputs.determinism_inputs())
    L = refs.full("hip")
    before = K.counters()["cpu_fallbacks"]
    with cpu_path():
        for f in gold["frames"]:
            src = data[f["input"]]
            for threads in (1, 4):
                frame = refs.kolmo_compress(L, src, f["level"], threads=threads)
                assert hashlib.sha256(frame.tobytes()).hexdigest() == f["sha256"], (f, threads)
            assert refs.kolmo_decompress(L, frame, len(src), threads=2).tobytes() == src
    assert K.counters()["cpu_fallbacks"] - before >= 2 * len(gold["frames"])
