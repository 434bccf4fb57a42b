"""Batches of independent blocks (kbwt_*_batch): every block's result equals the single-block result, which the
oracle and kolmo's own C pin (see test_gpu_parity.py). Covers empty and one-byte blocks, blocks that are not packed
(gaps between them), in-place device batches, and a mix of the measured generators.
"""
import sys

import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_blocks(seed, sizes):
    kinds = ["markov2", "mixed", "random", "dna", "repeats", "logs", "bf16", "csvcols"]
    return [K.generate(kinds[i % len(kinds)], seed + i, n, model_seed=seed) for i, n in enumerate(sizes)]


def check_batch(blocks, rs, tag):
    U, I = K.bwt_aux_batch(blocks, rs)
    for b, T in enumerate(blocks):
        eU, eI = refs.oracle_bwt(T, rs[b]) if T.size else (T, np.zeros(1, np.int32))
        assert np.array_equal(U[b], eU), f"{tag}: block {b} (n {T.size}) BWT differs"
        assert np.array_equal(I[b], eI), f"{tag}: block {b} (n {T.size}) indexes differ"
    back = K.unbwt_aux_batch(U, rs, I)
    for b, T in enumerate(blocks):
        assert np.array_equal(back[b], T), f"{tag}: block {b} inverse"


def test_batch_edge_sizes(gpu):
    sizes = [0, 1, 2, 3, 7, 100, 1000, 4095, 4096, 4097, 65536, 70001, 300000]
    blocks = mixed_blocks(100, sizes)
    rs = [max(2, 1 << K.codec18_rlog(max(n, 2))) for n in sizes]
    say("batch edge sizes", sizes)
    check_batch(blocks, rs, "edge sizes")


def test_batch_same_as_single(gpu):
    sizes = [5000, 123457, 40000, 2, 999999, 65537]
    blocks = mixed_blocks(200, sizes)
    rs = [1 << K.codec18_rlog(n) for n in sizes]
    say("batch vs single", sizes)
    U, I = K.bwt_aux_batch(blocks, rs)
    for b, T in enumerate(blocks):
        u1, i1 = K.bwt_aux(T, rs[b])
        assert np.array_equal(U[b], u1) and np.array_equal(I[b], i1), b


def test_batch_many_small(gpu):
    rng = np.random.default_rng(7)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(2), np.log(200000), 300))]
    blocks = mixed_blocks(300, sizes)
    rs = [1 << int(rng.integers(1, 12)) for _ in sizes]
    say("300 small blocks, total", sum(sizes))
    check_batch(blocks, rs, "many small")


def test_batch_host_inverse_repeated(gpu):
    """The host-pointer inverse decodes in place in a device buffer that comes back at the same address call after
    call; each call must read its own input, not lines of the previous call's output still held in a cache."""
    rng = np.random.default_rng(7)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(2), np.log(200000), 300))]
    blocks = mixed_blocks(300, sizes)
    rs = [1 << int(rng.integers(1, 12)) for _ in sizes]
    U, I = K.bwt_aux_batch(blocks, rs)
    say("host inverse repeated, total", sum(sizes))
    for rep in range(12):
        back = K.unbwt_aux_batch(U, rs, I)
        wrong = [b for b, T in enumerate(blocks) if not np.array_equal(back[b], T)]
        assert not wrong, f"call {rep}: blocks {wrong[:10]} differ"


def test_batch_not_a_bwt_block(gpu):
    """A block whose string is no BWT decodes to zeros over its whole length (the emission zeroes it tile by tile,
    about 150 tiles here), and the blocks around it in the same batch still decode exactly."""
    sizes = [50000, 300000, 7000, 123457]
    blocks = mixed_blocks(500, sizes)
    rs = [1 << K.codec18_rlog(n) for n in sizes]
    U, I = K.bwt_aux_batch(blocks, rs)
    U[1] = np.frombuffer(b"aabb" * 75000, np.uint8).copy()   # its LF map has more than one cycle
    say("batch with a block that is no BWT", sizes)
    back = K.unbwt_aux_batch(U, rs, I)
    assert back[1].size == sizes[1] and not back[1].any()
    for b in (0, 2, 3):
        assert np.array_equal(back[b], blocks[b]), b


def test_batch_device_gaps_and_in_place(gpu):
    import torch
    sizes = [70000, 3, 250000, 4096, 1, 33333]
    blocks = mixed_blocks(400, sizes)
    rs = [1 << K.codec18_rlog(max(n, 2)) for n in sizes]
    gaps = [17, 5, 0, 64, 3, 9]
    segs, total = K.make_segs(sizes, rs, gaps)
    buf = np.full(total + 1, 0xAB, np.uint8)
    for b, T in enumerate(blocks):
        buf[segs[b].offset: segs[b].offset + T.size] = T
    dT = torch.from_numpy(buf).cuda()
    dU = torch.full_like(dT, 0xCD)
    say("device batch with gaps", sizes, gaps)
    dU, I = K.dev_bwt_aux_batch(dT, segs, len(sizes), dU=dU)
    torch.cuda.synchronize()
    U = dU.cpu().numpy()
    k = 0
    for b, T in enumerate(blocks):
        eU, eI = K.bwt_aux(T, rs[b]) if T.size > 1 else (T, np.array([T.size], np.int32))
        o = segs[b].offset
        assert np.array_equal(U[o: o + T.size], eU), b
        ns = K.n_samples(T.size, rs[b])
        assert np.array_equal(I[k: k + ns], eI), b
        k += ns
    # bytes outside the blocks are untouched
    inside = np.zeros(total + 1, bool)
    for b in range(len(sizes)):
        inside[segs[b].offset: segs[b].offset + sizes[b]] = True
    assert (U[~inside] == 0xCD).all()
    # inverse in place over the device buffer
    K.dev_unbwt_aux_batch(dU, segs, len(sizes), I, dT=dU)
    torch.cuda.synchronize()
    back = dU.cpu().numpy()
    for b, T in enumerate(blocks):
        o = segs[b].offset
        assert np.array_equal(back[o: o + T.size], T), b


def test_batch_validation(gpu):
    segs, total = K.make_segs([100, 100], [4, 3])      # r = 3: not a power of two
    T = np.zeros(total + 1, np.uint8)
    U = np.zeros_like(T)
    I = np.zeros(64, np.int32)
    L = K.lib()
    assert L.kbwt_bwt_aux_batch(K._ptr(T), K._ptr(U), segs, 2, K._ptr(I, K._i32p)) == -1
    segs, total = K.make_segs([100, 100], [4, 4])
    segs[1].offset = 50                                   # overlapping blocks
    assert L.kbwt_bwt_aux_batch(K._ptr(T), K._ptr(U), segs, 2, K._ptr(I, K._i32p)) == -1
    segs, total = K.make_segs([10, 10], [10, 10])
    I = np.array([1, 11], np.int32)                       # second block's primary out of range
    assert L.kbwt_unbwt_aux_batch(K._ptr(T), K._ptr(U), segs, 2, K._ptr(I, K._i32p)) == -1


def test_batch_workload_sample(gpu):
    """A 64-block sample of the uneven-batch workload (sizes log-uniform 4 KiB-1 MiB, four generators)."""
    blocks, kinds, seeds = K.generate_batch(4, 64, 4096, 1 << 20, ["markov2", "mixed", "random", "dna"])
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]
    say("64-block workload sample, total", sum(b.size for b in blocks))
    check_batch(blocks, rs, "workload sample")


def test_gpu_batch_cfg4_full(gpu):
    """The whole batch the benchmark times (512 blocks of 4 KiB-16 MiB, 946 MiB, seed 4) in one batched forward and
    one batched inverse on device buffers: every block's BWT bytes and sampled indexes equal kolmo's own C
    (tests/golden/batch_cfg4.json, written by oracle/_ref/libkolmo_ref.so), and every block comes back."""
    import hashlib
    import json
    import os
    import torch
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_cfg4.json")))
    say("cfg4: generating", fx["count"], "blocks")
    blocks, _, _ = K.generate_batch(fx["seed"], fx["count"], fx["lo"], fx["hi"], fx["kinds"])
    sizes = [int(b.size) for b in blocks]
    assert sizes == [e["n"] for e in fx["blocks"]] and sum(sizes) == fx["total"]
    rs = [1 << K.codec18_rlog(n) for n in sizes]
    assert rs == [e["r"] for e in fx["blocks"]]
    segs, total = K.make_segs(sizes, rs)
    dT = torch.from_numpy(np.concatenate(blocks)).cuda()
    say("cfg4: batched forward of", total, "bytes")
    dU, I = K.dev_bwt_aux_batch(dT, segs, len(blocks))
    st = K.last_stats()
    say("cfg4: rounds", st["rounds"], "active", st["active"][:8])
    U = dU.cpu().numpy()
    k = 0
    for b, e in enumerate(fx["blocks"]):
        ns = K.n_samples(sizes[b], rs[b])
        o = segs[b].offset
        h = hashlib.sha256(U[o:o + sizes[b]].tobytes())
        h.update(np.asarray(I[k:k + ns], "<i4").tobytes())
        assert h.hexdigest() == e["sha_UI"], f"block {b} ({e['kind']}, n {sizes[b]}) differs from kolmo's"
        k += ns
    say("cfg4: batched inverse")
    back = K.dev_unbwt_aux_batch(dU, segs, len(blocks), I)
    torch.cuda.synchronize()
    assert torch.equal(back, dT), "batched inverse did not restore every block"
