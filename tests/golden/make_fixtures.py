"""Writes the committed fixtures of tests/golden/ from kolmo's own C, compiled from the kolmo tree by oracle/Makefile
(oracle/_ref/libkolmo_ref.so), and from the golden table of kolmo's tests/test_bwt_golden.h.

  fixtures.json    per input: the generator, seed and n, SHA-256 of the input, of the BWT bytes U and of the sampled
                   indexes I (int32 little-endian) for the codec-18 sampling interval, the primary index, and SHA-256
                   and size of the codec-18 stream (kolmo_bwt_compress with KOLMO_BWT_DEFAULTS).
  bsc_golden.json  the 63 fnv-1a hashes of libbsc 3.4.0 bsc_compress(LZP 15/72, BWT, static QLFC) streams over the
                   input kinds of tests/test_bwt.c:24-45, as the table holds them (data, not code).
  batch_cfg4.json  the 512-block batch the benchmark times (seed 4, sizes log-uniform 4 KiB-16 MiB, Markov text,
                   text + CSV, uniform bytes and DNA): per block its kind, n, r, SHA-256 of the block, and SHA-256 of
                   its BWT bytes followed by its sampled indexes (int32 little-endian), from kolmo's own C.

Run from the repository root after `make -C oracle`:  python tests/golden/make_fixtures.py [--batch-only]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import kolmo_amd as K  # noqa: E402
import refs  # noqa: E402

REF_TREE = "/root/reference"

INPUTS = [
    # name, generator, seed, n
    ("cfg1_markov2_1MiB", "markov2", 1, 1 << 20),
    ("cfg2_mixed_32MiB", "mixed", 2, 32 << 20),
    ("cfg3_repeats_128MiB", "repeats", 3, 128 << 20),
    ("mixed_1MiB", "mixed", 11, 1 << 20),
    ("dna_4MiB", "dna", 12, 4 << 20),
    ("random_1MiB", "random", 13, 1 << 20),
    ("repeats_8MiB", "repeats", 14, 8 << 20),
    ("logs_2MiB", "logs", 15, 2 << 20),
    ("bf16_2MiB", "bf16", 16, 2 << 20),
    ("csvcols_2MiB", "csvcols", 17, 2 << 20),
]


def entry(name, gen, seed, n):
    T = K.generate(gen, seed, n)
    rlog = K.codec18_rlog(n)
    r = 1 << rlog
    rc, U, I = refs.sais_bwt(refs.ref(), T, r, threads=8)
    assert rc == 0, rc
    rc2, S = refs.bwt18_compress(refs.ref(), T, threads=8)
    return {
        "name": name, "gen": gen, "seed": seed, "n": n, "r": r,
        "sha_T": refs.sha256(T), "sha_U": refs.sha256(U), "sha_I": refs.sha256(I.astype("<i4")),
        "primary": int(I[0]), "I": [int(x) for x in I],
        "codec18_size": int(rc2), "sha_codec18": refs.sha256(S) if rc2 > 0 else None,
    }


BATCH_CFG4 = {"seed": 4, "count": 512, "lo": 4096, "hi": 16 << 20, "kinds": ["markov2", "mixed", "random", "dna"]}


def batch_fixture():
    import hashlib
    from concurrent.futures import ThreadPoolExecutor
    b = BATCH_CFG4
    blocks, kind_of, _ = K.generate_batch(b["seed"], b["count"], b["lo"], b["hi"], b["kinds"])
    R = refs.ref()

    def one(k):
        T = blocks[k]
        r = 1 << K.codec18_rlog(T.size)
        rc, U, I = refs.sais_bwt(R, T, r, threads=2)
        assert rc == 0, rc
        h = hashlib.sha256(U.tobytes())
        h.update(I.astype("<i4").tobytes())
        return {"kind": {v: n for n, v in K.GEN_KINDS.items()}[int(kind_of[k])],
                "n": int(T.size), "r": r, "sha_T": refs.sha256(T), "sha_UI": h.hexdigest()}

    with ThreadPoolExecutor(max_workers=4) as ex:
        out = list(ex.map(one, range(len(blocks))))
    return dict(b, source="oracle/_ref/libkolmo_ref.so (kolmo 0.0.2 sais.c, built from source)",
                total=int(sum(x["n"] for x in out)), blocks=out)


def bsc_golden():
    text = open(os.path.join(REF_TREE, "tests", "test_bwt_golden.h")).read()
    rows = re.findall(r"\{([^}]*)\}", text)
    table = [[int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]+)ull", row)] for row in rows]
    assert len(table) == 9 and all(len(r) == 7 for r in table)
    sizes = [29, 1000, 65536, 70000, 262144, 1048576, 3000000]   # tests/test_bwt.c:47
    return {"sizes": sizes, "lzp_hash": 15, "lzp_min": 72, "coder": 1,
            "fnv1a": [[f"{v:016x}" for v in row] for row in table]}


def main():
    assert refs.have_ref(), "build oracle/_ref first (make -C oracle)"
    with open(os.path.join(HERE, "batch_cfg4.json"), "w") as f:
        json.dump(batch_fixture(), f, indent=0)
    if "--batch-only" in sys.argv:
        return
    out = [entry(*x) for x in INPUTS]
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump({"source": "oracle/_ref/libkolmo_ref.so (kolmo 0.0.2 sais.c + bwt_stream.c, built from source)",
                   "inputs": out}, f, indent=1)
    with open(os.path.join(HERE, "bsc_golden.json"), "w") as f:
        json.dump(bsc_golden(), f, indent=1)
    print("wrote", len(out), "fixtures")


if __name__ == "__main__":
    main()
