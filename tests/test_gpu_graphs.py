"""Launch sequences replayed from executable graphs give the bytes of direct launches. The forward's first sort and
the inverse are captured when a shape comes back and replayed from then on, two graphs each; the calls here alternate
between shapes so that a graph is replayed after calls of other shapes, evicted, and captured again."""
import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu


def test_gpu_graph_replays_alternating_shapes(gpu):
    import torch
    shapes = [("mixed", 11, 300000), ("markov2", 12, 1 << 20), ("dna", 13, 200000)]
    data = []
    for kind, seed, n in shapes:
        T = K.generate(kind, seed, n)
        r = 1 << K.codec18_rlog(n)
        eU, eI = refs.oracle_bwt(T, r)
        dT = torch.from_numpy(np.concatenate([T, np.zeros(1, np.uint8)])).cuda()
        dU, dB = torch.empty_like(dT), torch.empty_like(dT)
        segs, _ = K.make_segs([n], [r])
        data.append((T, eU, eI, dT, dU, dB, segs))
    c0 = K.counters()["graph_replays"]
    # 0 and 1 alternate (each keeps its graphs), 2 comes in and evicts the older, 0 comes back after it
    order = [0, 0, 1, 1, 0, 1, 0, 1, 2, 2, 2, 0, 0, 1, 2, 0]
    for step, i in enumerate(order):
        T, eU, eI, dT, dU, dB, segs = data[i]
        _, I = K.dev_bwt_aux_batch(dT, segs, 1, dU=dU)
        torch.cuda.synchronize()
        assert np.array_equal(dU[:T.size].cpu().numpy(), eU) and np.array_equal(I, eI), ("forward", step, i)
        K.dev_unbwt_aux_batch(dU, segs, 1, I, dT=dB)
        torch.cuda.synchronize()
        assert np.array_equal(dB[:T.size].cpu().numpy(), T), ("inverse", step, i)
    replays = K.counters()["graph_replays"] - c0
    print("graph launches", replays, flush=True)
    # every call after a shape's second one runs from a graph: forward and inverse, at least 2 x 8 of the 16 calls
    assert replays >= 16
