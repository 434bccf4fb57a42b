"""Whole kolmo frames through the drop-in (GPU).

kolmo's library built from its own sources with sais.c replaced by libkbwt_hip.so (oracle/Makefile,
oracle/_ref/libkolmo_full_hip.so) writes the frames kolmo publishes for levels 8 and 9
(benchmark/determinism/expected.sha256), at one and four threads, and decodes them. On larger generated inputs cut
into several blocks (concurrent calls from kolmo's block workers), its frames equal those of the unchanged library
(libkolmo_full_ref.so) byte for byte and decode through it; the library's own launch counters show that the
forward and inverse BWT of every block ran on the device.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import container_inputs
import kolmo_amd as K
import refs

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (refs.have_full("hip") and refs.have_full("ref")),
                                 reason="oracle/_ref/libkolmo_full_*.so not built (kolmo tree absent at build time)")]
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "container_golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def launches():
    L = K.lib()
    recs = (K.KernelTime * 64)()
    cnt = L.kbwt_profile_read(recs, 64)
    return {recs[i].name.decode(): recs[i].launches for i in range(cnt)}


def test_gpu_container_published_digests(gpu):
    H = refs.full("hip")
    data = dict(container_inputs.determinism_inputs())
    for f in GOLD["frames"]:
        src = data[f["input"]]
        for threads in (1, 4):
            say("frame", f["input"], "level", f["level"], "threads", threads)
            frame = refs.kolmo_compress(H, src, f["level"], threads=threads)
            assert hashlib.sha256(frame.tobytes()).hexdigest() == f["sha256"], (f, threads)
        back = refs.kolmo_decompress(H, frame, len(src), threads=2)
        assert back.tobytes() == src


@pytest.mark.parametrize("kind,seed,n,level", [("mixed", 11, 6 << 20, 8), ("markov2", 12, 5 << 20, 9),
                                               ("dna", 13, 3 << 20, 9), ("logs", 14, 4 << 20, 8)])
def test_gpu_container_multiblock_equals_reference(gpu, kind, seed, n, level):
    H, R = refs.full("hip"), refs.full("ref")
    src = K.generate(kind, seed, n)
    L = K.lib()
    L.kbwt_profile(1)
    try:
        say("multiblock", kind, n, "level", level)
        got = refs.kolmo_compress(H, src, level, threads=4, block=1 << 20)
        back = refs.kolmo_decompress(H, got, n, threads=4)
        counts = launches()
    finally:
        L.kbwt_profile(0)
    want = refs.kolmo_compress(R, src, level, threads=4, block=1 << 20)
    assert got.size == want.size and np.array_equal(got, want), "frame differs from kolmo's own build"
    assert np.array_equal(back, src), "frame does not decode through the drop-in"
    say("  device launches", {k: counts.get(k, 0) for k in ("bwt", "lf_pass")})
    # a frame of this kind and level sorts blocks of it: each forward ends with one k_bwt, each inverse builds LF once
    assert counts.get("bwt", 0) >= 1 and counts.get("lf_pass", 0) >= 1


def test_gpu_v12_frames_decode_to_the_originals(gpu):
    """kolmo's frozen-format corpus (tests/v12): every frame that holds block-sorted blocks (34 of the 357, written by
    kolmo at its max, -8, -9 and ultra settings) decodes through the drop-in to the original's SHA-256 at one and four
    threads, and each of its inverse BWTs ran on the device (the library's combined-call counter moves by the frame's
    inverse calls)."""
    V12 = json.load(open(os.path.join(HERE, "golden", "v12_frames.json")))
    H = refs.full("hip")
    for f in V12["frames"]:
        frame = np.frombuffer(open(os.path.join(HERE, "golden", "v12", f["file"]), "rb").read(), np.uint8)
        for threads in (1, 4):
            say("v12", f["corpus_path"], "threads", threads)
            c0 = K.counters()["combined_blocks"]
            back = refs.kolmo_decompress(H, frame, f["n"], threads=threads)
            assert hashlib.sha256(back.tobytes()).hexdigest() == f["sha256"], (f["file"], threads)
            assert K.counters()["combined_blocks"] - c0 == f["inverse_calls"], f["file"]
