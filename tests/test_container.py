"""Whole kolmo frames (CPU): the full library built from kolmo's sources by oracle/Makefile reproduces the frame
digests kolmo publishes for levels 8 and 9 (benchmark/determinism/expected.sha256), unchanged and with sais.c
replaced by the oracle. This pins the build recipe that the GPU container test (test_gpu_container.py) relinks
against libkbwt_hip.so, and shows that these frames go through the block-sorting entry points (the oracle's call
counter moves), so a digest match there is a statement about the GPU path.
"""
import hashlib
import json
import os

import pytest

import container_inputs
import refs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "container_golden.json")))


def inputs():
    return dict(container_inputs.determinism_inputs())


def test_determinism_inputs_reproduce():
    for name, data in inputs().items():
        e = GOLD["inputs"][name]
        assert len(data) == e["n"] and hashlib.sha256(data).hexdigest() == e["sha256"], name


@pytest.mark.parametrize("which", ["ref", "oracle"])
def test_full_library_frames_match_published_digests(which):
    if not refs.have_full(which):
        pytest.skip("oracle/_ref/libkolmo_full_*.so not built (kolmo tree absent)")
    L = refs.full(which)
    data = inputs()
    before = L.oracle_shim_calls(0) if which == "oracle" else 0
    for f in GOLD["frames"]:
        src = data[f["input"]]
        for threads in (1, 4):
            frame = refs.kolmo_compress(L, src, f["level"], threads=threads)
            assert hashlib.sha256(frame.tobytes()).hexdigest() == f["sha256"], (f, threads)
        back = refs.kolmo_decompress(L, frame, len(src), threads=2)
        assert back.tobytes() == src
    if which == "oracle":
        # every one of these frames has block-sorted blocks: the shim answered forward and inverse calls
        assert L.oracle_shim_calls(0) - before >= len(GOLD["frames"]) * 2
        assert L.oracle_shim_calls(1) > 0


V12 = json.load(open(os.path.join(HERE, "golden", "v12_frames.json")))


def v12_frame(f):
    return open(os.path.join(HERE, "golden", "v12", f["file"]), "rb").read()


def test_v12_frames_are_the_corpus_data():
    assert len(V12["frames"]) == 34
    for f in V12["frames"]:
        assert hashlib.sha256(v12_frame(f)).hexdigest() == f["frame_sha256"], f["file"]


@pytest.mark.parametrize("which", ["ref", "oracle"])
def test_v12_frames_decode_to_the_originals(which):
    """The frames of kolmo's frozen corpus (tests/v12) that hold block-sorted blocks decode to the originals' digests
    through kolmo's full library, unchanged and with sais.c replaced by the oracle (whose inverse runs for each)."""
    if not refs.have_full(which):
        pytest.skip("oracle/_ref/libkolmo_full_*.so not built (kolmo tree absent)")
    import numpy as np
    L = refs.full(which)
    for f in V12["frames"]:
        before = L.oracle_shim_calls(1) if which == "oracle" else 0
        back = refs.kolmo_decompress(L, np.frombuffer(v12_frame(f), np.uint8), f["n"], threads=2)
        assert hashlib.sha256(back.tobytes()).hexdigest() == f["sha256"], f["file"]
        if which == "oracle":
            assert L.oracle_shim_calls(1) - before == f["inverse_calls"], f["file"]
