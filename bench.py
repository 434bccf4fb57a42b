#!/usr/bin/env python3
"""Forward + inverse BWT throughput of kolmo's block-sorting path on one MI355X.

A step is one forward BWT (suffix sort, BWT bytes, primary and sampled indexes) plus one inverse BWT of the workload,
already resident in HBM, through the library's device-resident entry points. Output: one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload cfg2|cfg1|cfg3|cfg4] [--no-cpu] [--no-extras]

Workloads (sampling interval r by codec 18's rule for every block):
  cfg2 (default)  one 32 MiB block (seed 2): 70% order-3 Markov text over 96 symbols (Dirichlet(0.2) rows) and 30%
                  CSV lines of 8 int32 random-walk columns
  cfg1            one 1 MiB block (seed 1): order-2 Markov over 64 symbols (Dirichlet(0.3) rows)
  cfg3            one 128 MiB deep-repeat block (seed 3)
  cfg4            512 independent blocks (seed 4), sizes log-uniform 4 KiB-16 MiB, each from the cfg1 or cfg2
                  generators, uniform bytes or DNA; batched calls, timed against one call per block

Besides throughput it reports, for the dominant kernel, the algorithmic bytes per launch over its average duration
measured with HIP events on its own stream inside the timed region (roofline), and kolmo's own CPU path
(oracle/_ref, compiled from kolmo's sources) timed on this host's cores on a bounded sample of the same workload.
The default cfg2 line also carries, under config: the 128 MiB deep-repeat block ("deep_repeats"), the batched form
on cfg4 against one call per block ("batched"),
the shipped router on the held-out router blocks ("router"), and the cfg2 block through the host-pointer entry
points, PCIe included ("pcie_inclusive"), and the same block as 1 MiB blocks through those entry points from 16
concurrent callers as kolmo's block workers make them, against one caller, from Python threads and from C threads
("concurrent_calls"); --no-extras leaves them out.
"""
import argparse
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
    "regroup": 18.0,       # read key 8 + value 4, write ISA 4, SA row byte 1 (+ the group lists: <= 4 per row)
    "round_keys": 32.0,    # read pos, grp, SA, ISA[i+h] (4 each), write key 8 + value 4, plus the active row's SA read
    "bwt": 2.0,            # read the row byte, write the BWT byte
    "lf_pass": 5.0,        # read L 1, write LF 4
    "lf_pairs": 17.0,      # read LF 4 and L 1, gather LF(LF) 4, write the pair 8
    "walk_count": 5.0,     # one LF read and one staged byte per step
    "walk_emit": 2.0,      # one staged byte read and written per step
    "rank_jump": 24.0,     # read two links, write one
    "fill_blk": 2.0,
    "scan_groups": 8.0,    # read a group size, write its first-row offset
    "large_elems": 8.0,    # write pos + group head per row of a large group
    "seg_gather": 12.0,    # read the row's suffix 4 and ISA[i+h] 4, write the sort key 4
    "seg_sort": 18.0,      # read key 4 + suffix 4, write SA 4, ISA 4, row byte 1, read it 1 (T[i-1] or carried)
    "isa_part": 16.0,      # read and write one (suffix, head) pair
    "isa_scatter": 12.0,   # read a pair, write ISA
    "emit_compact": 0.0,   # packs the next round's group list (a few bytes per group, not per row)
}

# The same with every random gather into a table far larger than an XCD's L2 charged one 128-byte line, the read
# request size of the GPU's L2 (rocprofv3: TCC_EA0_RDREQ counts 128 bytes per request for these kernels): the counting
# walk's pair loads (one per two steps), the doubling rounds' ISA[i + h] gathers. This line model is what measured HBM
# traffic is checked against; the algorithmic one is what roofline.achieved counts.
LINE_BYTES = dict(KERNEL_BYTES, walk_count=1.0 + 128.0 / 2, seg_gather=12.0 - 4 + 128, round_keys=32.0 - 4 + 128)

WORKLOADS = {
    "cfg1": ("markov2", 1, 1 << 20, "one 1 MiB block (seed 1): order-2 Markov over 64 symbols, Dirichlet(0.3)"),
    "cfg2": ("mixed", 2, 32 << 20, "one 32 MiB block (seed 2): 70% order-3 Markov text (96 symbols, Dirichlet(0.2)) "
             "interleaved with 30% CSV lines of 8 int32 random-walk columns (N(0,50) steps)"),
    "cfg3": ("repeats", 3, 128 << 20, "one 128 MiB block (seed 3): 1 MiB random, then copies of earlier substrings "
             "(64 B-4 MiB, 0.01% substitutions) and one 2 MiB single-byte run"),
}
BATCH = {"cfg4": (4, 512, 4096, 16 << 20, ["markov2", "mixed", "random", "dna"],
                  "512 blocks (seed 4), sizes log-uniform 4 KiB-16 MiB, from the order-2 Markov and text+CSV "
                  "generators, uniform bytes and DNA with N and tandem repeats")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) and scaled
    by the same counters on a plain copy, as recorded in profiles/traffic_<workload>_*.md ('| kernel | ... | bytes per
    launch |'); None when no profile of that workload holds the kernel."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for f in sorted(os.listdir(pdir)):
        if not (f.startswith(f"traffic_{workload}_") and f.endswith(".md")):
            continue
        for line in open(os.path.join(pdir, f)):
            m = re.match(r"\|\s*(\S+)\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$", line)
            if m and m.group(1) == kernel:
                try:
                    return float(m.group(2))
                except ValueError:
                    pass
    return None


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(n, 64))


def cpu_time(R, blocks, rs, threads, budget_s, reps):
    """kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux of kolmo's own C (oracle/_ref, built from kolmo's sources) over the
    blocks in order, each best of `reps` after one warm-up, until about budget_s seconds of work; every round trip
    checked. Returns (bytes, seconds, forward seconds, inverse seconds, blocks used)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refs
    done_bytes, done_s, fwd_s, inv_s, used = 0, 0.0, 0.0, 0.0, 0
    for T, r in zip(blocks, rs):
        best = None
        for rep in range(reps + 1):
            t0 = time.perf_counter()
            rc, U, I = refs.sais_bwt(R, T, r, threads=threads)
            t1 = time.perf_counter()
            rc2, back = refs.sais_unbwt(R, U, r, I, threads=threads)
            t2 = time.perf_counter()
            assert rc == 0 and rc2 == 0 and np.array_equal(back, T)
            if rep > 0 and (best is None or t2 - t0 < best[0]):
                best = (t2 - t0, t1 - t0, t2 - t1)
        done_bytes += T.size
        done_s += best[0]
        fwd_s += best[1]
        inv_s += best[2]
        used += 1
        if done_s * (reps + 1) >= budget_s:
            break
    return done_bytes, done_s, fwd_s, inv_s, used


def cpu_baseline(blocks, rs, budget_s=12.0, reps=3, cfg1=True):
    """kolmo's own path on this host's cores, as the measurement plan asks: threads = nproc (the reported value) and
    threads = 1, best of 3 after one warm-up, with the CPU model; beside it the 1 MiB cfg1 block at both thread counts
    and kolmo's whole codec-18 stage (kolmo_bwt_compress / kolmo_bwt_decompress with its defaults) at nproc."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refs
    if not refs.have_ref():
        return None
    R = refs.ref()
    nproc = host_threads()
    runs = {}
    for th in (nproc, 1):
        b, sec, f, i, used = cpu_time(R, blocks, rs, th, budget_s, reps)
        runs[th] = {"value": round(b / sec / 2**20, 2), "fwd_ms": round(f * 1e3, 1), "inv_ms": round(i * 1e3, 1),
                    "bytes": int(b), "blocks": used}
    what = (f"the full {blocks[0].size >> 20} MiB block" if len(blocks) == 1 else
            f"the first {runs[nproc]['blocks']} of {len(blocks)} blocks ({runs[nproc]['bytes'] / 2**20:.1f} MiB)")
    out = {"value": runs[nproc]["value"], "unit": "MiB/s", "cores": nproc, "kind": "reference",
           "sample": f"{what}, kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux (kolmo 0.0.2 sais.c, gcc -O3 "
                     f"-march=x86-64-v3) at threads={nproc}, best of {reps} after a warm-up; fwd "
                     f"{runs[nproc]['fwd_ms']:.0f} ms, inv {runs[nproc]['inv_ms']:.0f} ms",
           "cpu_model": cpu_model(), "threads_nproc": runs[nproc], "threads_1": runs[1]}
    if cfg1:
        import kolmo_amd as K
        gen, seed, n, _ = WORKLOADS["cfg1"]
        T1 = K.generate(gen, seed, n)
        r1 = 1 << K.codec18_rlog(n)
        c = {}
        for th in (nproc, 1):
            b, sec, f, i, _ = cpu_time(R, [T1], [r1], th, budget_s, reps)
            c[f"threads_{'nproc' if th == nproc else 1}"] = {"value": round(b / sec / 2**20, 2),
                                                             "fwd_ms": round(f * 1e3, 2), "inv_ms": round(i * 1e3, 2)}
        out["cfg1"] = dict(c, workload="cfg1: " + WORKLOADS["cfg1"][3])
    # the whole codec-18 stage (LZP, BWT, QLFC) for context, on the first block
    T = blocks[0]
    best_c = best_d = None
    for _ in range(2):
        t0 = time.perf_counter()
        rc, S = refs.bwt18_compress(R, T, threads=nproc)
        t1 = time.perf_counter()
        rc2, back = refs.bwt18_decompress(R, S, T.size, threads=nproc)
        t2 = time.perf_counter()
        assert rc > 0 and rc2 == 0 and back.tobytes() == T.tobytes()
        best_c = t1 - t0 if best_c is None else min(best_c, t1 - t0)
        best_d = t2 - t1 if best_d is None else min(best_d, t2 - t1)
    out["codec18_stage"] = {"compress": round(T.size / best_c / 2**20, 2), "decompress": round(T.size / best_d / 2**20, 2),
                            "unit": "MiB/s", "threads": nproc,
                            "what": "kolmo_bwt_compress / kolmo_bwt_decompress with KOLMO_BWT_DEFAULTS, best of 2"}
    return out


def traffic_total_from_profiles(workload):
    """Measured HBM bytes per step of the whole path on `workload`, from the 'bytes per step, all kernels' line of
    profiles/traffic_<workload>_*.md (rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy); None when no
    such profile exists."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for f in sorted(os.listdir(pdir)):
        if f.startswith(f"traffic_{workload}_") and f.endswith(".md"):
            for line in open(os.path.join(pdir, f)):
                m = re.match(r"Bytes per step, all kernels: ([0-9.eE+]+)", line)
                if m:
                    return float(m.group(1))
    return None


def roofline_from(kernels, steps, ms_step, dom, d, workload):
    """Roofline of kernel `dom` from its record `d` (summed ms, launches, elements) taken in the timed region; the
    per-kernel table of the whole path comes from the separate all-kernel pass `kernels`."""
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = KERNEL_BYTES.get(dom, 0.0) * d["elems"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom, workload), "kernel": dom,
                "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": int(bytes_per_launch)}
    path_bytes = sum(KERNEL_BYTES.get(k, 0.0) * v["elems"] for k, v in kernels.items()) / steps
    line_bytes = sum(LINE_BYTES.get(k, 0.0) * v["elems"] for k, v in kernels.items()) / steps
    measured = traffic_total_from_profiles(workload)
    path = {"bytes_per_step": int(path_bytes), "achieved_GBs": round(path_bytes / (ms_step * 1e-3) / 1e9, 1),
            "frac": round(path_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "line_model_bytes_per_step": int(line_bytes),
            "measured_traffic_per_step": measured,
            "measured_over_line_model": round(measured / line_bytes, 3) if measured else None,
            "kernel_ms_per_step": {k: round(v["ms"] / steps, 4) for k, v in sorted(kernels.items(),
                                   key=lambda kv: -kv[1]["ms"])},
            "kernel_bytes_per_step": {k: int(KERNEL_BYTES.get(k, 0.0) * v["elems"] / steps) for k, v in kernels.items()},
            "kernel_line_bytes_per_step": {k: int(LINE_BYTES.get(k, 0.0) * v["elems"] / steps)
                                           for k, v in kernels.items()}}
    return roofline, path, dom, achieved


def model_bytes(n, st):
    """Algorithmic bytes of one forward + inverse from the doubling profile (m_r = active[r] suffixes sorted in round r
    with p_r radix passes): B_fwd = 13n + sum_r (8 + 24 p_r + 16) m_r + sum_{r>=1} 20 m_r + 6n, B_inv = 16n. Rounds
    after the first sort groups of up to 512 rows inside one workgroup, counted as one pass."""
    act, passes = st["active"], st["passes"]
    b = 13 * n + 6 * n + 16 * n
    for r, m in enumerate(act):
        p = passes[r] if r == 0 else max(1, passes[r])
        b += (8 + 24 * p + 16) * m + (20 * m if r >= 1 else 0)
    return b


def _profile_records(K):
    L = K.lib()
    recs = (K.KernelTime * 64)()
    cnt = L.kbwt_profile_read(recs, 64)
    L.kbwt_profile(0)
    return {recs[i].name.decode(): {"ms": recs[i].ms, "launches": recs[i].launches, "elems": recs[i].elems}
            for i in range(cnt)}


def timed_steps(K, steps, step):
    """Per-kernel pass, then the timed region.

    1. `steps` steps with HIP events around every launch: {kernel: {ms, launches, elems}}, and the dominant kernel
       (the most device time).
    2. The timed region: `steps` steps with HIP events around the dominant kernel's launches alone (on the stream they
       are launched on), so the region runs as it would unprofiled (launch graphs included) and still measures that
       kernel's launch durations live.

    Returns (seconds of the timed region, all-kernel table, dominant kernel, its record from the timed region)."""
    import torch
    L = K.lib()
    L.kbwt_profile(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    kernels = _profile_records(K)
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    log(f"bench: all-kernel pass {tp / steps * 1e3:.3f} ms/step wall, "
        f"{sum(v['ms'] for v in kernels.values()) / steps:.3f} ms/step in kernels; dominant {dom}")
    L.kbwt_profile_only(dom.encode())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    live = _profile_records(K)
    if dom not in live:
        raise SystemExit(f"no timed launches of {dom} in the timed region")
    return t1 - t0, kernels, dom, live[dom]


def calib_copy(K, nbytes, dev, stream):
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    src.fill_(7)
    torch.cuda.synchronize()
    rc = K.lib().kbwt_dev_copy(dst.data_ptr(), src.data_ptr(), nbytes, stream.cuda_stream)
    if rc != 0 or not torch.equal(src, dst):
        raise SystemExit("calibration copy failed")
    log(f"bench: calibration copy of {nbytes} bytes done")


def run_single(args, K):
    import hashlib
    import torch
    gen, seed, n, desc = WORKLOADS[args.workload]
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
    for _ in range(max(1, args.warmup)):
        _, I = K.dev_bwt_aux(dT, r, dU=dU, stream=stream)
        st = K.last_stats()                   # the forward's doubling profile (the inverse's call resets the stats)
        K.dev_unbwt_aux(dU, r, I, dT=dB, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(dB, dT):
        raise SystemExit("inverse BWT did not restore the block")
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "fixtures.json")))["inputs"]
    entry = next((e for e in fx if e["gen"] == gen and e["seed"] == seed and e["n"] == n), None)
    parity = None
    if entry is not None:
        parity = (hashlib.sha256(dU.cpu().numpy().tobytes()).hexdigest() == entry["sha_U"]
                  and [int(x) for x in I] == entry["I"])
        if not parity:
            raise SystemExit("BWT differs from kolmo's reference output")
    if args.calib_copy:
        calib_copy(K, args.calib_copy, dev, stream)

    def step():
        _, Ix = K.dev_bwt_aux(dT, r, dU=dU, stream=stream)
        K.dev_unbwt_aux(dU, r, Ix, dT=dB, stream=stream)

    log(f"bench: timing {args.steps} steps")
    elapsed, kernels, dom, drec = timed_steps(K, args.steps, step)
    ms_step = elapsed / args.steps * 1e3
    value = n * args.steps / elapsed / 2**20
    roofline, path, dom, achieved = roofline_from(kernels, args.steps, ms_step, dom, drec, args.workload)
    mb = model_bytes(n, st)
    path["model_bytes_per_step"] = int(mb)
    path["model_frac"] = round(mb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    log(f"bench: {value:.1f} MiB/s, {ms_step:.3f} ms/step, dominant {dom} {achieved:.0f} GB/s, "
        f"path {path['model_frac']:.3f} of HBM peak")
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (kolmo's own C on this host)")
        cpu = cpu_baseline([T], [r])
    config = {"workload": f"{args.workload}: {desc}", "n": n, "r": r, "bit_exact_vs_reference": parity,
              "rounds": st["rounds"], "active": st["active"], "sigma": st["sigma"], "h0": st["h0"]}
    return value, ms_step, config, roofline, path, cpu


def run_batch(args, K):
    import numpy as np
    import torch
    seed, count, lo, hi, kinds, desc = BATCH[args.workload]
    log(f"bench: generating {args.workload} ({count} blocks)")
    blocks, kind_of, seeds = K.generate_batch(seed, count, lo, hi, kinds)
    sizes = [int(b.size) for b in blocks]
    rs = [1 << K.codec18_rlog(n) for n in sizes]
    segs, total = K.make_segs(sizes, rs)
    host = np.concatenate(blocks)
    dev = torch.device("cuda:0")
    dT = torch.from_numpy(host).to(dev)
    dU = torch.empty_like(dT)
    dB = torch.empty_like(dT)
    stream = torch.cuda.current_stream()
    log("bench: warmup + round trip of every block")
    I = None
    for _ in range(max(1, args.warmup)):
        _, I = K.dev_bwt_aux_batch(dT, segs, count, dU=dU, stream=stream)
        st = K.last_stats()                   # the batched forward's doubling profile
        K.dev_unbwt_aux_batch(dU, segs, count, I, dT=dB, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(dB, dT):
        raise SystemExit("batched inverse BWT did not restore the blocks")
    if args.calib_copy:
        calib_copy(K, args.calib_copy, dev, stream)
    # the batched result of a few blocks against the single-block calls
    U = dU.cpu().numpy()
    k = 0
    for b in range(0 if getattr(args, "no_per_block", False) else count):
        ns = K.n_samples(sizes[b], rs[b])
        if b < 8:
            u1, i1 = K.bwt_aux(blocks[b], rs[b])
            o = segs[b].offset
            if not (np.array_equal(U[o: o + sizes[b]], u1) and np.array_equal(I[k: k + ns], i1)):
                raise SystemExit(f"batched BWT of block {b} differs from the single-block BWT")
        k += ns

    def step():
        _, Ix = K.dev_bwt_aux_batch(dT, segs, count, dU=dU, stream=stream)
        K.dev_unbwt_aux_batch(dU, segs, count, Ix, dT=dB, stream=stream)

    log(f"bench: timing {args.steps} batched steps")
    elapsed, kernels, dom, drec = timed_steps(K, args.steps, step)
    ms_step = elapsed / args.steps * 1e3
    value = total * args.steps / elapsed / 2**20

    views = [(dT[segs[b].offset: segs[b].offset + sizes[b]], dU[segs[b].offset: segs[b].offset + sizes[b]],
              dB[segs[b].offset: segs[b].offset + sizes[b]]) for b in range(count)]

    def step_per_block():
        for b in range(count):
            t, u, x = views[b]
            _, Ib = K.dev_bwt_aux(t, rs[b], dU=u, stream=stream)
            K.dev_unbwt_aux(u, rs[b], Ib, dT=x, stream=stream)

    per_block_s = per_block = None
    if not getattr(args, "no_per_block", False):
        log("bench: timing one call per block")
        step_per_block()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_per_block()
        torch.cuda.synchronize()
        per_block_s = time.perf_counter() - t0
        per_block = total / per_block_s / 2**20
    roofline, path, dom, achieved = roofline_from(kernels, args.steps, ms_step, dom, drec, args.workload)
    mb = model_bytes(total, st)
    path["model_bytes_per_step"] = int(mb)
    path["model_frac"] = round(mb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if per_block:
        log(f"bench: batched {value:.1f} MiB/s, one call per block {per_block:.1f} MiB/s ({value / per_block:.2f}x)")
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (kolmo's own C on this host)")
        cpu = cpu_baseline(blocks, rs, cfg1=False)
    config = {"workload": f"{args.workload}: {desc}", "blocks": count, "total_bytes": total,
              "rounds": st["rounds"], "active": st["active"],
              "one_call_per_block": {"value": round(per_block, 2), "unit": "MiB/s",
                                     "ms_per_step": round(per_block_s * 1e3, 3)} if per_block else None,
              "batched_over_per_block": round(value / per_block, 3) if per_block else None}
    return value, ms_step, config, roofline, path, cpu


def pcie_inclusive(K, reps=3):
    """The cfg2 block through the host-pointer entry points (copies over PCIe included), best of reps."""
    gen, seed, n, _ = WORKLOADS["cfg2"]
    T = K.generate(gen, seed, n)
    r = 1 << K.codec18_rlog(n)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        U, I = K.bwt_aux(T, r)
        K.unbwt_aux(U, r, I)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    log(f"bench: host-pointer forward + inverse {n / best / 2**20:.1f} MiB/s")
    return {"value": round(n / best / 2**20, 2), "unit": "MiB/s", "ms_per_step": round(best * 1e3, 3),
            "what": "kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux on host buffers, best of 3"}


def concurrent_calls(K, workers=16, reps=3):
    """kolmo's block workers on the host-pointer entry points: the cfg2 block cut into 1 MiB blocks, each worker
    taking blocks in turn and calling kolmo_sais_bwt_aux then kolmo_sais_unbwt_aux (PCIe included), against the same
    calls made one after another from one thread. Concurrent calls run as batches inside the library."""
    import threading
    import numpy as np
    gen, seed, n, _ = WORKLOADS["cfg2"]
    T = K.generate(gen, seed, n)
    blocks = [T[o:o + (1 << 20)].copy() for o in range(0, n, 1 << 20)]
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]

    def one(k):
        U, I = K.bwt_aux(blocks[k], rs[k])
        back = K.unbwt_aux(U, rs[k], I)
        if not np.array_equal(back, blocks[k]):
            raise SystemExit("round trip failed in the concurrent-call measurement")

    def serial():
        for k in range(len(blocks)):
            one(k)

    def parallel():
        ks = list(range(len(blocks)))
        lock = threading.Lock()

        def worker():
            while True:
                with lock:
                    if not ks:
                        return
                    k = ks.pop(0)
                one(k)

        ts = [threading.Thread(target=worker) for _ in range(workers)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()

    res = {}
    for name, f in (("one_thread", serial), ("workers", parallel)):
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res[name] = round(n / best / 2**20, 2)
    log(f"bench: 1 MiB blocks through kolmo_sais_*: {res['one_thread']} MiB/s from one thread, "
        f"{res['workers']} MiB/s from {workers} workers")
    # the same calls from C threads (libkbwt_workers.so), with no interpreter lock between them
    c1 = n / K.block_workers(T, 1 << 20, 1, reps) / 2**20
    cw = n / K.block_workers(T, 1 << 20, workers, reps) / 2**20
    log(f"bench: the same from C threads: {c1:.2f} MiB/s from one thread, {cw:.2f} MiB/s from {workers} workers")
    return {"blocks": len(blocks), "block_bytes": 1 << 20, "workers": workers, "unit": "MiB/s",
            "one_thread": res["one_thread"], "concurrent_workers": res["workers"],
            "c_threads": {"one_thread": round(c1, 2), "concurrent_workers": round(cw, 2)},
            "what": "kolmo_sais_bwt_aux + kolmo_sais_unbwt_aux per block on host buffers (PCIe included), best of 3; "
                    "from Python threads, and from C threads (c_threads)"}


def router_summary(sample=256):
    """The shipped router on the held-out fifth of the router workload (8192 blocks of 256 KiB-4 MiB, seed 5): its
    bytes against exhaustive best-per-block (every class coded, the smallest kept) and its coder throughput. The
    features of `sample` held-out blocks are recomputed live here (regenerated blocks, one batched forward with
    features, timed on the device stream) and must equal the committed ones; their device seconds per byte, scaled to
    the held-out bytes, are the features' share of the router's time. Coder times are the committed single-thread CPU
    times of kolmo's coders. Also the same figures on the out-of-distribution set (real files of kolmo's tests, from
    its committed features: rebuilding those files needs kolmo's own library)."""
    import numpy as np
    import torch
    import kolmo_amd as K
    from kolmo_amd import router as R
    lab_p = os.path.join(ROOT, "tests", "golden", "router_labels.npz")
    feat_p = os.path.join(ROOT, "tests", "golden", "router_features.npz")
    ood_p = os.path.join(ROOT, "tests", "golden", "router_ood.npz")
    if not (os.path.exists(lab_p) and os.path.exists(feat_p) and os.path.exists(R.MODEL)):
        return None
    lab = np.load(lab_p)
    F = np.load(feat_p)["features"]
    P, th = R.load_model()
    protos, radius = R.load_gate()
    sizes = R.outcome_sizes(lab["n"], lab["size_bwt"], lab["size_klzma"], lab["size_bwt2"])
    test = np.arange(len(F)) % 5 == 4
    tb, tk, t2 = (lab[k].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2"))
    # live features of a sample of held-out blocks
    idx = np.flatnonzero(test)[:sample]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=16) as ex:
        blocks = list(ex.map(lambda i: R.workload_block(lab["kind"][i], lab["seed"][i], lab["n"][i], model_seed=5),
                             idx))
    sizes_s = [int(b.size) for b in blocks]
    segs, total = K.make_segs(sizes_s, [1 << K.codec18_rlog(x) for x in sizes_s])
    dT = torch.from_numpy(np.concatenate(blocks + [np.zeros(1, np.uint8)])).cuda()
    stream = torch.cuda.current_stream()
    R.dev_features_batch(dT, segs, len(blocks), stream=stream)        # warm-up (workspace, code objects)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    _, _, Fl = R.dev_features_batch(dT, segs, len(blocks), stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t_live = e0.elapsed_time(e1) * 1e-3
    if np.max(np.abs(Fl - F[idx])) > 1e-6:
        raise SystemExit("live router features differ from the committed ones")
    share = t_live / total * float(lab["n"][test].sum())
    t0 = time.perf_counter()
    dec = R.decide(R.predict(F[test], P), th, None if protos is None else R.gate(F[test], protos), radius)
    t_pred = time.perf_counter() - t0
    ev = R.evaluate(dec, sizes[test], tb[test], tk[test], t2[test], t_features=[share + t_pred])
    exh = R.evaluate(np.full(test.sum(), -1), sizes[test], tb[test], tk[test], t2[test])
    log(f"bench: router {ev['loss_pct']:+.3f}% bytes at {exh['seconds'] / ev['seconds']:.2f}x exhaustive throughput "
        f"(features {total / t_live / 2**20:.0f} MiB/s live)")
    out = {"blocks": int(test.sum()), "size_loss_pct": round(ev["loss_pct"], 4),
           "throughput_over_exhaustive": round(exh["seconds"] / ev["seconds"], 3),
           "fallback_blocks": ev["fallback_blocks"], "classes": "block sorting, KLZMA, store, block sorting of halves",
           "gate": None if protos is None else {"prototypes": int(len(protos)), "radius": round(radius, 4)},
           "features_live": {"blocks": len(blocks), "bytes": int(total), "device_seconds": round(t_live, 5),
                             "MiB_per_s": round(total / t_live / 2**20, 1)},
           "what": "held-out fifth of 8192 blocks; coder times are single-thread CPU times of kolmo's coders; the "
                   "features' share is timed live on a sample"}
    if os.path.exists(ood_p):
        d = np.load(ood_p)
        if "features" in d:
            so = R.outcome_sizes(d["n"], d["size_bwt"], d["size_klzma"], d["size_bwt2"])
            to = [d[k].astype(np.float64) for k in ("t_bwt", "t_klzma", "t_bwt2")]
            go = None if protos is None else R.gate(d["features"], protos)
            eo = R.evaluate(R.decide(R.predict(d["features"], P), th, go, radius), so, *to)
            xo = R.evaluate(np.full(len(so), -1), so, *to)
            bo = R.evaluate(np.full(len(so), R.BWT), so, *to)
            out["out_of_distribution"] = {"blocks": int(len(so)), "size_loss_pct": round(eo["loss_pct"], 4),
                                          "always_bwt_loss_pct": round(bo["loss_pct"], 4),
                                          "throughput_over_exhaustive": round(xo["seconds"] / eo["seconds"], 3),
                                          "fallback_blocks": eo["fallback_blocks"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2", choices=sorted(list(WORKLOADS) + list(BATCH)))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run one plain device copy of this many bytes (the reference for byte counters)")
    ap.add_argument("--no-extras", action="store_true",
                    help="cfg2 only: skip the deep-repeat and batched lines, the router summary, the PCIe-inclusive "
                         "rate and the concurrent calls")
    ap.add_argument("--no-per-block", action="store_true",
                    help="cfg4 only: skip the single-block calls (the one-call-per-block comparison and the check of "
                         "8 blocks against them), for profiled runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): blocks are independent and a multi-GPU "
              "run is not part of this build", file=sys.stderr)
        sys.exit(2)
    import kolmo_amd as K
    if not K.device_ok():
        raise SystemExit("no usable gfx950 device (or libkbwt_hip.so missing)")
    run = run_batch if args.workload in BATCH else run_single
    value, ms_step, config, roofline, path, cpu = run(args, K)
    if args.workload == "cfg2" and not args.no_extras:
        config["pcie_inclusive"] = pcie_inclusive(K)
        config["concurrent_calls"] = concurrent_calls(K)
        dargs = argparse.Namespace(**vars(args))
        dargs.workload, dargs.steps, dargs.warmup, dargs.no_cpu, dargs.calib_copy = "cfg3", 2, 1, True, 0
        dv, dms, dcfg, droof, dpath, _ = run_single(dargs, K)
        config["deep_repeats"] = {"workload": dcfg["workload"], "value": round(dv, 2), "unit": "MiB/s",
                                  "ms_per_step": round(dms, 3), "rounds": dcfg["rounds"],
                                  "bit_exact_vs_reference": dcfg["bit_exact_vs_reference"],
                                  "dominant_kernel": droof["kernel"], "dominant_frac": droof["frac"],
                                  "model_frac": dpath["model_frac"]}
        bargs = argparse.Namespace(**vars(args))
        bargs.workload, bargs.steps, bargs.warmup, bargs.no_cpu = "cfg4", 2, 1, True
        bv, bms, bcfg, _, _, _ = run_batch(bargs, K)
        config["batched"] = {"workload": bcfg["workload"], "value": round(bv, 2), "unit": "MiB/s",
                             "ms_per_step": round(bms, 3), "one_call_per_block": bcfg["one_call_per_block"],
                             "batched_over_per_block": bcfg["batched_over_per_block"]}
        config["router"] = router_summary()
    out = {
        "metric": "fwd+inv BWT MiB/s on 1 MI355X: batched vs plain GPU vs host CPU; router size loss %", "value": round(value, 2), "unit": "MiB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic", "config": config,
        "roofline": roofline, "path_roofline": path, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
