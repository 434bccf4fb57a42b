"""Copies the frames of kolmo's frozen-format corpus (tests/v12, MANIFEST: frame and original) that hold
block-sorted blocks into tests/golden/v12/, and writes tests/golden/v12_frames.json: per frame its path in the
corpus, the original's path, size and SHA-256, and the frame's SHA-256. The frames are data kolmo's own tests hold;
the originals are not copied, only their digests.

A frame holds block-sorted blocks when decoding it through kolmo's full library relinked over the oracle
(oracle/_ref/libkolmo_full_oracle.so) calls the inverse BWT: the oracle shim counts those calls.

Run from the repository root after `make -C oracle`:  python tests/golden/make_v12_frames.py
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refs  # noqa: E402

CORPUS = "/root/reference/tests/v12"


def main():
    L = refs.full("oracle")
    rows = [line.split() for line in open(os.path.join(CORPUS, "MANIFEST")) if line.strip()]
    out_dir = os.path.join(HERE, "v12")
    os.makedirs(out_dir, exist_ok=True)
    frames = []
    for frame_path, orig_path in rows:
        orig = open(os.path.join(CORPUS, orig_path), "rb").read()
        frame = open(os.path.join(CORPUS, frame_path), "rb").read()
        before = L.oracle_shim_calls(1)
        back = refs.kolmo_decompress(L, np.frombuffer(frame, np.uint8), len(orig))
        assert back.tobytes() == orig, frame_path
        calls = L.oracle_shim_calls(1) - before
        if not calls:
            continue
        name = frame_path.replace("/", "__")
        shutil.copyfile(os.path.join(CORPUS, frame_path), os.path.join(out_dir, name))
        frames.append({"file": name, "corpus_path": frame_path, "original": orig_path, "n": len(orig),
                       "sha256": hashlib.sha256(orig).hexdigest(), "frame_sha256": hashlib.sha256(frame).hexdigest(),
                       "inverse_calls": int(calls)})
    with open(os.path.join(HERE, "v12_frames.json"), "w") as f:
        json.dump({"source": "kolmo 0.0.2 tests/v12 (MANIFEST): the frames whose decode runs the inverse BWT",
                   "frames": frames}, f, indent=1)
    print("copied", len(frames), "frames of", len(rows))


if __name__ == "__main__":
    main()
