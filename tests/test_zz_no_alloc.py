"""kbwt_dev_unbwt_aux_noalloc (GPU): the inverse that allocates nothing gives kbwt_dev_unbwt_aux's output.

Called through the C ABI on device buffers: two calls in a row, the second on a smaller block (so the static scratch
holds the first call's contents), then an in-place call, an input and output that are not 16-byte aligned, a string
that is no BWT, the largest block it takes and the first one it refuses. Every output is compared byte for byte with
the allocating entry point's on the same input (and with the original block, which the round trip must restore).
"""
import ctypes
import sys

import numpy as np
import pytest

import kolmo_amd as K

pytestmark = pytest.mark.gpu
MAX_N = 1 << 22


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(dU, r, I, dT, stream):
    I = np.ascontiguousarray(I, np.int32)
    return K.lib().kbwt_dev_unbwt_aux_noalloc(ctypes.c_void_p(dU.data_ptr()), ctypes.c_void_p(dT.data_ptr()),
                                               dU.numel(), r, K._ptr(I, K._i32p), ctypes.c_void_p(stream.cuda_stream))


def forward(T, dev, stream):
    import torch
    r = 1 << K.codec18_rlog(T.size)
    dT = torch.from_numpy(T.copy()).to(dev)
    dU, I = K.dev_bwt_aux(dT, r, stream=stream)
    return dU, r, I


def test_noalloc_matches_the_allocating_inverse(gpu):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    cases = [("mixed", 2, 3 << 20), ("markov2", 1, 200000), ("repeats", 3, MAX_N), ("dna", 4, 77777),
             ("random", 5, 4099)]
    for gen, seed, n in cases:
        say("noalloc", gen, seed, n)
        T = K.generate(gen, seed, n)
        dU, r, I = forward(T, dev, stream)
        want = K.dev_unbwt_aux(dU, r, I, stream=stream)
        got = torch.full_like(dU, 0xAB)
        assert noalloc(dU, r, I, got, stream) == 0, K.lib().kbwt_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, want), (gen, n)
        assert np.array_equal(got.cpu().numpy(), T), (gen, n)
        # r = n: only the primary index
        got1 = torch.zeros_like(dU)
        assert noalloc(dU, n, I[:1], got1, stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(got1, want)


def test_noalloc_in_place_unaligned_and_not_a_bwt(gpu):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    T = K.generate("mixed", 9, 1 << 20)
    dU, r, I = forward(T, dev, stream)
    want = K.dev_unbwt_aux(dU, r, I, stream=stream)
    say("noalloc in place")
    io = dU.clone()
    assert noalloc(io, r, I, io, stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(io, want)
    say("noalloc unaligned input and output")
    big = torch.zeros(T.size + 64, dtype=torch.uint8, device=dev)
    big[3:3 + T.size] = dU
    out = torch.zeros(T.size + 64, dtype=torch.uint8, device=dev)
    assert noalloc(big[3:3 + T.size], r, I, out[5:5 + T.size], stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(out[5:5 + T.size], want)
    assert not torch.any(out[:5]) and not torch.any(out[5 + T.size:])
    say("noalloc not a bwt")
    bad = dU.clone()
    bad[10], bad[20000] = dU[20000].item(), dU[10].item()
    want_bad = K.dev_unbwt_aux(bad, r, I, stream=stream)
    got_bad = torch.full_like(bad, 7)
    assert noalloc(bad, r, I, got_bad, stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(got_bad, want_bad)


def test_noalloc_limits(gpu):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    say("noalloc limits")
    d = torch.zeros(MAX_N + 1, dtype=torch.uint8, device=dev)
    r = 1 << K.codec18_rlog(MAX_N + 1)
    I = np.ones(K.n_samples(MAX_N + 1, r), np.int32)
    assert noalloc(d, r, I, d, stream) == K.KBWT_ERR_ARG
    assert "4194304" in K.lib().kbwt_last_error().decode()
    one = torch.tensor([42], dtype=torch.uint8, device=dev)
    o = torch.zeros_like(one)
    assert noalloc(one, 1, np.array([1], np.int32), o, stream) == 0
    torch.cuda.synchronize()
    assert o.item() == 42
    # and the allocating entry point is unchanged after the static one ran
    T = K.generate("markov2", 1, 1 << 20)
    dU, r, I = forward(T, dev, stream)
    back = K.dev_unbwt_aux(dU, r, I, stream=stream)
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy(), T)
