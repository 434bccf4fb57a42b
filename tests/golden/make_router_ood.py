#!/usr/bin/env python3
"""Labels of the router's out-of-distribution set: real files kolmo's own tests hold, which none of the router's six
generators produced. Each file is one block:

  - the originals of the frames of tests/golden/v12/ (kolmo's frozen corpus, tests/v12: executables, images, audio,
    CSV and JSON files, FASTQ and FASTA, logs, text), rebuilt by decoding the frames with kolmo's own library
    (oracle/_ref/libkolmo_full_ref.so), so no file has to be copied;
  - the three inputs of kolmo's determinism check that tests/container_inputs.py regenerates (text.txt, table.csv,
    log.txt).

For every block it records, as make_router_labels.py does for the generated blocks, the payload size and
single-thread time of kolmo's coders from oracle/_ref: kolmo_bwt_compress with KOLMO_BWT_DEFAULTS on the block and on
its two halves, and kolmo_klzma_compress at native level 8. Output tests/golden/router_ood.npz (no pickles): name,
n, sha256, size_bwt, size_klzma, size_bwt2 (-1: stored), t_bwt, t_klzma, t_bwt2.

The second mode, on a machine with the GPU, adds the router's features of every block (one batched forward,
kbwt_dev_bwt_aux_batch_features) as `features` and writes the result to OUT:

    python tests/golden/make_router_ood.py
    python tests/golden/make_router_ood.py --features OUT
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))


def ood_blocks():
    """[(name, uint8 array)]: the distinct originals of the committed v12 frames, then the determinism inputs."""
    import container_inputs
    import refs
    L = refs.full("ref")
    meta = json.load(open(os.path.join(HERE, "v12_frames.json")))["frames"]
    seen, out = set(), []
    for f in meta:
        if f["sha256"] in seen:
            continue
        seen.add(f["sha256"])
        frame = np.frombuffer(open(os.path.join(HERE, "v12", f["file"]), "rb").read(), np.uint8)
        data = refs.kolmo_decompress(L, frame, f["n"], threads=2)
        assert hashlib.sha256(data.tobytes()).hexdigest() == f["sha256"]
        out.append((f["original"], data.copy()))
    for name, data in container_inputs.determinism_inputs():
        out.append(("determinism/" + name, np.frombuffer(data, np.uint8).copy()))
    return out


def main():
    import ctypes
    import refs
    import make_router_labels as M
    M._init()
    L = M._lib
    rows = []
    for name, T in ood_blocks():
        t0 = time.perf_counter()
        rc, _ = refs.bwt18_compress(L, T, threads=1)
        t1 = time.perf_counter()
        out = np.empty(T.size + (T.size >> 2) + 4096, np.uint8)
        k = L.kolmo_klzma_compress(T.ctypes.data, T.size, out.ctypes.data, out.size, 8)
        t2 = time.perf_counter()
        s2, t_2 = M.bwt_halves(T)
        rows.append((name, T.size, hashlib.sha256(T.tobytes()).hexdigest(), rc if rc > 0 else -1,
                     int(k) if k > 0 else -1, s2, t1 - t0, t2 - t1, t_2))
        print(f"{name}: n {T.size} bwt {rows[-1][3]} klzma {rows[-1][4]} halves {s2}", flush=True)
    cols = list(zip(*rows))
    np.savez_compressed(os.path.join(HERE, "router_ood.npz"), name=np.array(cols[0]), n=np.array(cols[1], np.int64),
                        sha256=np.array(cols[2]), size_bwt=np.array(cols[3], np.int64),
                        size_klzma=np.array(cols[4], np.int64), size_bwt2=np.array(cols[5], np.int64),
                        t_bwt=np.array(cols[6], np.float32), t_klzma=np.array(cols[7], np.float32),
                        t_bwt2=np.array(cols[8], np.float32))
    print("wrote", len(rows), "blocks")


def main_features(out):
    sys.path.insert(0, ROOT)
    from kolmo_amd import router as R
    d = dict(np.load(os.path.join(HERE, "router_ood.npz")))
    blocks = ood_blocks()
    assert [b[0] for b in blocks] == list(d["name"])
    d["features"] = R.features_of([b[1] for b in blocks]).astype(np.float32)
    np.savez_compressed(out, **d)
    print("wrote features of", len(blocks), "blocks to", out)


if __name__ == "__main__":
    if "--features" in sys.argv:
        main_features(sys.argv[sys.argv.index("--features") + 1])
    else:
        main()
