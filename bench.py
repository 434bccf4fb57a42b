#!/usr/bin/env python3
"""Forward + inverse BWT throughput of kolmo's block-sorting path on one MI355X.

A step is one forward BWT (suffix sort, BWT bytes, primary and sampled indexes) plus one inverse BWT of one block
already resident in HBM, through the library's device-resident entry points (kbwt_dev_bwt_aux / kbwt_dev_unbwt_aux).
Default workload: one 32 MiB block (seed 2) of 70% order-3 Markov text over 96 symbols (Dirichlet(0.2) rows) and 30%
CSV lines of 8 int32 random-walk columns, sampling interval r = 2^22 (codec 18's rule). Output: one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME] [--no-cpu]

Besides the throughput it reports, for the dominant kernel, the algorithmic bytes per launch over its average
duration measured with HIP events on its own stream inside the timed region (roofline), and kolmo's own CPU path
(oracle/_ref, compiled from kolmo's sources) timed on this host's cores on the same block (cpu_baseline).
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# Algorithmic bytes per element of each kernel (reads + writes the algorithm needs, counted once).
KERNEL_BYTES = {
    "sort_pass": 24.0,     # read key 8 + value 4, write key 8 + value 4
    "sort_pass_T": 13.0,   # read T 1, write key 8 + value 4
    "hist_T": 1.0,         # read T
    "byte_hist": 1.0,
    "hist_keys": 8.0,      # read key
    "regroup": 16.0,       # read key 8 + value 4, write ISA 4 (compaction writes of the few active rows not counted)
    "round_keys": 32.0,    # read pos, grp, SA, ISA[i+h] (4 each), write key 8 + value 4, plus the active row's SA read
    "bwt": 6.0,            # read SA 4 + T 1, write U 1
    "lf_pass": 5.0,        # read L 1, write LF 4
    "walk_count": 4.0,     # one LF read per step
    "walk_emit": 5.0,      # one LF read and one byte write per step
    "rank_jump": 24.0,     # read two links, write one
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workloads():
    import kolmo_amd as K
    return {
        "cfg2": ("mixed", 2, 32 << 20, "one 32 MiB block (seed 2): 70% order-3 Markov text (96 symbols, Dirichlet(0.2)) "
                 "interleaved with 30% CSV lines of 8 int32 random-walk columns (N(0,50) steps)"),
        "cfg1": ("markov2", 1, 1 << 20, "one 1 MiB block (seed 1): order-2 Markov over 64 symbols, Dirichlet(0.3)"),
        "cfg3": ("repeats", 3, 128 << 20, "one 128 MiB block (seed 3): 1 MiB random, then copies of earlier substrings "
                 "(64 B-4 MiB, 0.01% substitutions) and one 2 MiB single-byte run"),
    }, K


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) and scaled by the same
    counters on a plain copy, as recorded in profiles/traffic_*.md ('| kernel | ... | bytes per launch |')."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for f in sorted(os.listdir(pdir)):
        if not (f.startswith("traffic_") and f.endswith(".md")):
            continue
        for line in open(os.path.join(pdir, f)):
            m = re.match(r"\|\s*(\S+)\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$", line)
            if m and m.group(1) == kernel:
                try:
                    return float(m.group(2))
                except ValueError:
                    pass
    return None


def cpu_baseline(T, r):
    """kolmo's own kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux (oracle/_ref, built from kolmo's C), on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import refs
    if not refs.have_ref():
        return None
    R = refs.ref()
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    n = T.size
    best = None
    for rep in range(2):
        t0 = time.perf_counter()
        rc, U, I = refs.sais_bwt(R, T, r, threads=threads)
        t1 = time.perf_counter()
        rc2, back = refs.sais_unbwt(R, U, r, I, threads=threads)
        t2 = time.perf_counter()
        assert rc == 0 and rc2 == 0 and np.array_equal(back, T)
        if best is None or t2 - t0 < best[0]:
            best = (t2 - t0, t1 - t0, t2 - t1)
    return {"value": round(n / best[0] / 2**20, 2), "unit": "MiB/s", "cores": threads, "kind": "reference",
            "sample": f"the full {n >> 20} MiB block, kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux (kolmo 0.0.2 sais.c, "
                      f"gcc -O3 -march=x86-64-v3) at threads={threads}, best of 2; fwd {best[1]*1e3:.0f} ms, "
                      f"inv {best[2]*1e3:.0f} ms"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run one plain device copy of this many bytes (the reference for byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): blocks are independent and a multi-GPU "
              "run is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    W, K = workloads()
    gen, seed, n, desc = W[args.workload]
    if not K.device_ok():
        raise SystemExit("no usable gfx950 device (or libkbwt_hip.so missing)")
    log(f"bench: generating {args.workload} ({n} bytes)")
    T = K.generate(gen, seed, n)
    r = 1 << K.codec18_rlog(n)
    dev = torch.device("cuda:0")
    dT = torch.from_numpy(T.copy()).to(dev)
    dU = torch.empty_like(dT)
    dB = torch.empty_like(dT)
    stream = torch.cuda.current_stream()

    log("bench: warmup + parity check")
    I = None
    for w in range(max(1, args.warmup)):
        _, I = K.dev_bwt_aux(dT, r, dU=dU, stream=stream)
        K.dev_unbwt_aux(dU, r, I, dT=dB, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(dB, dT):
        raise SystemExit("inverse BWT did not restore the block")
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "fixtures.json")))["inputs"]
    ref_entry = next((e for e in fx if e["gen"] == gen and e["seed"] == seed and e["n"] == n), None)
    parity = None
    if ref_entry is not None:
        import hashlib
        parity = (hashlib.sha256(dU.cpu().numpy().tobytes()).hexdigest() == ref_entry["sha_U"]
                  and [int(x) for x in I] == ref_entry["I"])
        if not parity:
            raise SystemExit("BWT differs from kolmo's reference output")
    st = K.last_stats()
    if args.calib_copy:
        src = torch.empty(args.calib_copy, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        src.fill_(7)
        torch.cuda.synchronize()
        rc = K.lib().kbwt_dev_copy(dst.data_ptr(), src.data_ptr(), args.calib_copy, stream.cuda_stream)
        if rc != 0 or not torch.equal(src, dst):
            raise SystemExit("calibration copy failed")
        log(f"bench: calibration copy of {args.calib_copy} bytes done")

    log(f"bench: timing {args.steps} steps")
    L = K.lib()
    L.kbwt_profile(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, I = K.dev_bwt_aux(dT, r, dU=dU, stream=stream)
        K.dev_unbwt_aux(dU, r, I, dT=dB, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    recs = (K.KernelTime * 64)()
    cnt = L.kbwt_profile_read(recs, 64)
    L.kbwt_profile(0)
    elapsed = t1 - t0
    ms_step = elapsed / args.steps * 1e3
    value = n * args.steps / elapsed / 2**20

    kernels = {}
    for i in range(cnt):
        rec = recs[i]
        kernels[rec.name.decode()] = {"ms": rec.ms, "launches": rec.launches, "elems": rec.elems}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = KERNEL_BYTES.get(dom, 0.0) * d["elems"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": int(bytes_per_launch)}
    # whole-path byte model: sum over kernels of algorithmic bytes, over the measured step time
    path_bytes = sum(KERNEL_BYTES.get(k, 0.0) * v["elems"] for k, v in kernels.items()) / args.steps
    path = {"bytes_per_step": int(path_bytes), "achieved_GBs": round(path_bytes / (ms_step * 1e-3) / 1e9, 1),
            "frac": round(path_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_ms_per_step": {k: round(v["ms"] / args.steps, 4) for k, v in sorted(kernels.items(),
                                   key=lambda kv: -kv[1]["ms"])}}
    log(f"bench: {value:.1f} MiB/s, {ms_step:.3f} ms/step, dominant {dom} {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (kolmo's own C on this host)")
        cpu = cpu_baseline(T, r)

    out = {
        "metric": "fwd+inv BWT MiB/s on 1 MI355X", "value": round(value, 2), "unit": "MiB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "n": n, "r": r, "bit_exact_vs_reference": parity,
                   "rounds": st["rounds"], "active": st["active"], "sigma": st["sigma"], "h0": st["h0"]},
        "roofline": roofline, "path_roofline": path, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
