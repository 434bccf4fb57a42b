"""The autotuner (kbwt_autotune): it picks launch parameters per size class by timing the caller's blocks, the
pick is at least as fast as where it started (within timing noise), and no parameter changes any output: results
under the tuned and under every individual parameter value equal the oracle's."""
import numpy as np
import pytest

import kolmo_amd as K
import refs

pytestmark = pytest.mark.gpu


def test_gpu_autotune_picks_and_keeps_parity(gpu):
    import torch
    T = K.generate("mixed", 77, 3 << 20)
    r = 1 << K.codec18_rlog(T.size)
    segs, total = K.make_segs([T.size], [r])
    dT = torch.from_numpy(np.concatenate([T, np.zeros(1, np.uint8)])).cuda()
    cls = int(np.ceil(np.log2(T.size)))
    try:
        best, ms, ms0 = K.autotune(dT, segs, 1)
        print("autotune", best, ms, "from", ms0, flush=True)
        assert ms <= ms0 * 1.02
        assert K.get_tuning(cls) == best
        eU, eI = refs.oracle_bwt(T, r)
        U, I = K.bwt_aux(T, r)
        assert np.array_equal(U, eU) and np.array_equal(I, eI)
        assert np.array_equal(K.unbwt_aux(U, r, I), T)
    finally:
        K.reset_tuning()


@pytest.mark.parametrize("t", [(1, 4, 0), (2, 5, 8), (3, 7, 64), (4, 6, 1024), (5, 3, 16), (6, 8, 2)])
def test_gpu_every_setting_is_exact(gpu, t):
    try:
        K.set_tuning(t)
        for kind, n in [("mixed", 700000), ("repeats", 500000), ("dna", 300000)]:
            T = K.generate(kind, 78, n)
            r = 1 << K.codec18_rlog(n)
            U, I = K.bwt_aux(T, r)
            eU, eI = refs.oracle_bwt(T, r)
            assert np.array_equal(U, eU) and np.array_equal(I, eI), (t, kind)
            assert np.array_equal(K.unbwt_aux(U, r, I), T), (t, kind)
    finally:
        K.reset_tuning()
