"""Parity of the HIP path with the oracle and with kolmo's own C (GPU).

Bar: bit-exact. The BWT bytes, the primary index and every sampled index equal the reference's; the inverse returns
the original bytes; codec-5 and codec-18 streams written through the drop-in relink equal kolmo's own bytes.
Inputs follow kolmo's tests (tests/test_gpu_sa.c:74-98 edge plan and random inputs, tests/test_bwt.c kinds and
sizes, the golden hashes of tests/test_bwt_golden.h) plus the measured workloads against committed fixtures.
"""
import json
import os
import sys

import numpy as np
import pytest

import kolmo_amd as K
import refs
from test_oracle import edge_inputs, random_inputs

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_case(T, r, tag):
    U, I = K.bwt_aux(T, r)
    eU, eI = refs.oracle_bwt(T, r)
    assert np.array_equal(U, eU), f"{tag}: BWT bytes differ"
    assert np.array_equal(I, eI), f"{tag}: sampled indexes differ"
    back = K.unbwt_aux(U, r, I)
    assert np.array_equal(back, T), f"{tag}: inverse"
    back1 = K.unbwt_aux(U, T.size, I[:1])
    assert np.array_equal(back1, T), f"{tag}: inverse with r = n"


def test_gpu_tiny(gpu):
    for s in [b"ab", b"ba", b"aa", b"abracadabra", b"mississippi", b"\x00\x00\x00", b"banana" * 7]:
        T = np.frombuffer(s, np.uint8).copy()
        say("tiny", s[:12], len(s))
        check_case(T, 2, s)


def test_gpu_edge_plan(gpu):
    say("edge plan n = 2..299 x 7 kinds")
    for kind, n, T, r in edge_inputs():
        check_case(T, r, f"edge kind {kind} n {n}")


def test_gpu_random_inputs(gpu):
    for it, T in random_inputs(count=60, big=3000000, seed=5):
        say("random", it, T.size)
        check_case(T, 1 << K.codec18_rlog(T.size), f"random {it} n {T.size}")


def test_gpu_sampling_intervals(gpu):
    """Every r from 2 up to past n, on one input (I has (n-1)/r + 1 entries)."""
    T = K.generate("mixed", 21, 40000)
    for rl in range(1, 17):
        say("r = 2 **", rl)
        check_case(T, 1 << rl, f"r {1 << rl}")


def test_gpu_bsc_kinds(gpu):
    """tests/test_bwt.c kinds x sizes, against the oracle."""
    for kind in range(9):
        for n in [2, 3, 7, 100, 4096, 65536, 70001, 300000, 1000003]:
            say("bsc kind", kind, "n", n)
            T = K.generate_bsc_kind(kind, n)
            check_case(T, 1 << K.codec18_rlog(n), f"kind {kind} n {n}")


@pytest.mark.skipif(not refs.have_ref(), reason="oracle/_ref not built")
def test_gpu_codec5_golden_hashes(gpu):
    """kolmo's codec-5 glue relinked against libkbwt_hip.so writes libbsc 3.4.0's bytes (tests/test_bwt_golden.h),
    and decodes them back."""
    g = json.load(open(os.path.join(HERE, "golden", "bsc_golden.json")))
    C = refs.codec("hip")
    for kind in range(9):
        for j, n in enumerate(g["sizes"]):
            say("golden kind", kind, "n", n)
            T = K.generate_bsc_kind(kind, n)
            out = refs.bsc_compress(C, T, g["lzp_hash"], g["lzp_min"], g["coder"])
            assert f"{refs.fnv1a(out):016x}" == g["fnv1a"][kind][j], (kind, n)
            back = np.empty(n + 1, np.uint8)
            rc = C.kolmo_bsc_decompress(refs._p(out), out.size, refs._p(back), n, 1)
            assert rc == 0 and np.array_equal(back[:n], T)


@pytest.mark.skipif(not refs.have_ref(), reason="oracle/_ref not built")
def test_gpu_codec18_identical(gpu):
    """codec-18 streams (kolmo_bwt_compress, KOLMO_BWT_DEFAULTS and three other parameter sets of tests/test_bwt.c:128)
    are byte-identical with and without the drop-in, and decode through it."""
    R, C = refs.ref(), refs.codec("hip")
    params = [refs.BWT_DEFAULTS, (15, 72, 1, 0, 0, 0), (0, 0, 2, 3, 0, 0), (16, 32, 1, 12, 1 << 16, 0)]
    for kind in range(9):
        for n in [2, 3, 40, 1000, 65536, 300000, 1048576]:
            T = K.generate_bsc_kind(kind, n)
            for q, p in enumerate(params):
                say("codec18 kind", kind, "n", n, "params", q)
                ra, a = refs.bwt18_compress(R, T, p)
                rb, b = refs.bwt18_compress(C, T, p)
                assert ra == rb and np.array_equal(a, b), (kind, n, q, ra, rb)
                if ra > 0:
                    rc, back = refs.bwt18_decompress(C, b, n)
                    assert rc == 0 and np.array_equal(back, T)


def fixtures():
    return json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["inputs"]


@pytest.mark.parametrize("e", fixtures(), ids=lambda e: e["name"])
def test_gpu_fixtures(gpu, e):
    """The measured workloads and the other generators at full size: SHA-256 of U and I equal the reference's."""
    say("fixture", e["name"], e["n"])
    T = K.generate(e["gen"], e["seed"], e["n"])
    assert refs.sha256(T) == e["sha_T"]
    U, I = K.bwt_aux(T, e["r"])
    st = K.last_stats()
    say("  stats", {k: st[k] for k in ("sigma", "h0", "rounds", "ms_fwd")}, "active", st["active"][:6])
    assert refs.sha256(U) == e["sha_U"], "BWT bytes differ"
    assert [int(x) for x in I] == e["I"], "sampled indexes differ"
    back = K.unbwt_aux(U, e["r"], I)
    assert np.array_equal(back, T), "inverse"


def test_gpu_device_resident_api(gpu):
    """kbwt_dev_* on tensors already in HBM give the same bytes as the host-pointer entry points, also in place."""
    import torch
    T = K.generate("mixed", 31, 3 << 20)
    r = 1 << K.codec18_rlog(T.size)
    U, I = K.bwt_aux(T, r)
    dT = torch.from_numpy(T.copy()).cuda()
    dU, dI = K.dev_bwt_aux(dT, r)
    torch.cuda.synchronize()
    assert np.array_equal(dU.cpu().numpy(), U) and np.array_equal(dI, I)
    back = K.dev_unbwt_aux(dU, r, dI)
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy(), T)
    dX = torch.from_numpy(T.copy()).cuda()
    K.dev_bwt_aux(dX, r, dU=dX)                     # in place
    assert np.array_equal(dX.cpu().numpy(), U)
    K.dev_unbwt_aux(dX, r, I, dT=dX)                # in place
    assert np.array_equal(dX.cpu().numpy(), T)


def test_gpu_isa_window_edges(gpu):
    """Single blocks around the edges of the first round's ISA assembly: 32 K-suffix workgroup ranges that end inside
    or exactly at a block's end, partition windows narrower than, as wide as and twice or four times as wide as a
    workgroup's range (blocks up to 8 MiB, to 16 MiB, to 32 MiB), and sizes one off those edges."""
    for kind, n in [("mixed", 32767), ("mixed", 32768), ("mixed", 32769), ("dna", 98305), ("logs", 131071),
                    ("csvcols", 131073), ("mixed", (1 << 23) - 1), ("markov2", (1 << 23) + 1),
                    ("repeats", (1 << 24) + 12345)]:
        T = K.generate(kind, 77, n)
        say("isa window edges", kind, n)
        check_case(T, 1 << K.codec18_rlog(n), f"isa edges {kind} n {n}")


def test_gpu_not_a_bwt(gpu):
    """A byte string that is no BWT (several LF cycles) decodes to defined bytes with code 0, as kolmo's decoder
    returns 0 there too; the caller's checksum rejects it."""
    U = np.frombuffer(b"aabb" * 1000, np.uint8).copy()   # the LF map of this string has more than one cycle
    back = K.unbwt_aux(U, U.size, np.array([1], np.int32))
    assert back.size == U.size
    back2 = K.unbwt_aux(U, U.size, np.array([1], np.int32))
    assert np.array_equal(back, back2)
    assert not back.any()                                  # every byte zero


@pytest.mark.parametrize("rank_max", ["0", "2", "1024"])
def test_gpu_round_paths(gpu, rank_max, monkeypatch):
    """Both ways a doubling round splits small groups (radix sort in LDS, direct ranking), forced one way or the
    other through KBWT_SEG_RANK, on inputs with many rounds of small and mid-sized groups."""
    monkeypatch.setenv("KBWT_SEG_RANK", rank_max)
    for kind, n in [("repeats", 1 << 20), ("dna", 700000), ("logs", 300000), ("csvcols", 200000), ("mixed", 500000)]:
        T = K.generate(kind, 31, n)
        say("round paths", rank_max, kind, n)
        check_case(T, 1 << K.codec18_rlog(n), f"{kind} n {n} rank_max {rank_max}")


def test_gpu_inverse_graph_replays(gpu):
    """Repeated inverses of one shape run from a captured HIP graph (the second call with a shape captures it, later
    calls replay it); every replay decodes its own input, also with forward sorts and other shapes in between that
    reuse the lookback status words and the workspace."""
    n = 3 << 19
    r = 1 << K.codec18_rlog(n)
    for i, kind in enumerate(["mixed", "dna", "repeats", "logs", "mixed", "csvcols"]):
        T = K.generate(kind, 40 + i, n)
        say("graph replay", i, kind, n)
        U, I = K.bwt_aux(T, r)
        assert np.array_equal(K.unbwt_aux(U, r, I), T), f"replay {i} {kind}"
        if i == 3:                                   # another shape, then back to the captured one
            T2 = K.generate("dna", 7, 100000)
            r2 = 1 << K.codec18_rlog(T2.size)
            U2, I2 = K.bwt_aux(T2, r2)
            assert np.array_equal(K.unbwt_aux(U2, r2, I2), T2)
