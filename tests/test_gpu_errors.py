"""The error path on the device (GPU): faults injected from the host, recovery, per-call isolation, the CPU path.

Every fault here is injected on the host (kbwt_fault_inject): the fault point fails before the HIP runtime is
called, so no kernel is launched differently and every launch queued before the fault completes normally. What is
checked is what kolmo's callers see (src/codecs/bwt.c:30-37 turn a non-zero code into "code this block another way"
or "corrupt"): the code of the failing call, that nobody else's call fails or waits for ever, that the next call is
exact, and, with the CPU path enabled, that kolmo's frames stay byte-identical (src/arch/metal/cubwt_metal.c:4-7).
"""
import hashlib
import json
import os
import sys
import threading

import numpy as np
import pytest

import kolmo_amd as K
import container_inputs
import refs

pytestmark = [pytest.mark.gpu, pytest.mark.fallback]
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def code_of(f, *a):
    try:
        f(*a)
        return 0
    except K.KbwtError as e:
        return e.code


def test_gpu_fault_at_every_point_then_exact(gpu):
    """For each fault point of one forward and one inverse call (every launch check, stream synchronisation and
    host copy, in order, and the first allocations of a fresh workspace), inject a fault there: the call returns -3
    (-2 for an allocation), and the next call, with nothing injected, is exact. The block is big enough for the
    doubling rounds, the pinned staging path (>= 1 MiB) and the inverse's graph replay."""
    T = K.generate("repeats", 11, 3 << 20)
    r = 1 << K.codec18_rlog(T.size)
    eU, eI = refs.oracle_bwt(T, r)
    K.set_cpu_fallback(False)
    U, I = K.bwt_aux(T, r)
    assert np.array_equal(U, eU) and np.array_equal(I, eI)
    fired_total = 0
    for site in ("launch", "sync", "copy", "alloc"):
        for direction in ("fwd", "inv"):
            k = 1
            while True:
                if site == "alloc":
                    if k > 6:
                        break
                    K.lib().kbwt_release()             # a fresh workspace allocates on its next call
                say("fault", site, k, direction)
                K.fault_inject(f"{site}:{k}")
                before = K.counters()["injected_faults"]
                rc = code_of(K.bwt_aux, T, r) if direction == "fwd" else code_of(K.unbwt_aux, eU, r, eI)
                K.fault_inject(None)
                fired = K.counters()["injected_faults"] - before
                if fired == 0:
                    assert rc == 0
                    break                                  # past the call's last fault point of this kind
                fired_total += 1
                assert rc == (K.KBWT_ERR_NOMEM if site == "alloc" else K.KBWT_ERR_DEVICE), (site, k, direction, rc)
                assert "injected" in K.lib().kbwt_last_error().decode()
                if direction == "fwd":
                    U, I = K.bwt_aux(T, r)
                    assert np.array_equal(U, eU) and np.array_equal(I, eI), (site, k)
                else:
                    assert np.array_equal(K.unbwt_aux(eU, r, eI), T), (site, k)
                k += 1
    say("faults injected:", fired_total)
    assert fired_total > 40


def test_gpu_one_fault_among_concurrent_callers(gpu):
    """16 callers mix forward and inverse calls; one launch fault is injected while they run. Every caller returns;
    at most the one call the fault hit alone fails (a combined batch that fails runs its blocks again one by one), and
    every other result is exact."""
    K.set_cpu_fallback(False)
    blocks = [K.generate("mixed", 40 + k, 200000 + 7919 * k) for k in range(16)]
    rs = [1 << K.codec18_rlog(b.size) for b in blocks]
    ref = [refs.oracle_bwt(b, r) for b, r in zip(blocks, rs)]
    for trial, k in enumerate((2, 9, 25, 60)):
        say("trial", trial, "fault at launch", k)
        codes, bad, lock = [], [], threading.Lock()
        c0 = K.counters()
        start = threading.Barrier(16)

        def worker(j):
            start.wait()
            try:
                if j % 2:
                    U, I = K.bwt_aux(blocks[j], rs[j])
                    ok = np.array_equal(U, ref[j][0]) and np.array_equal(I, ref[j][1])
                else:
                    ok = np.array_equal(K.unbwt_aux(ref[j][0], rs[j], ref[j][1]), blocks[j])
                rc = 0
            except K.KbwtError as e:
                rc, ok = e.code, True
            with lock:
                codes.append(rc)
                if not ok:
                    bad.append(j)

        K.fault_inject(f"launch:{k}")
        ts = [threading.Thread(target=worker, args=(j,)) for j in range(16)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=120)
        K.fault_inject(None)
        assert not any(t.is_alive() for t in ts), "a caller is still waiting"
        c1 = K.counters()
        say("codes", sorted(codes), "counters", {x: c1[x] - c0[x] for x in c1})
        assert not bad, bad
        assert len(codes) == 16 and set(codes) <= {0, K.KBWT_ERR_DEVICE}
        assert codes.count(K.KBWT_ERR_DEVICE) <= c1["injected_faults"] - c0["injected_faults"] <= 1


@pytest.mark.skipif(not refs.have_full("hip"), reason="oracle/_ref/libkolmo_full_hip.so not built (kolmo tree absent)")
def test_gpu_frames_stay_identical_with_the_cpu_path(gpu):
    """kolmo's full library relinked over libkbwt_hip.so, CPU path on: with a fault injected now and then, and with
    every device call failing, the level-8/9 frames equal kolmo's published digests and decode."""
    gold = json.load(open(os.path.join(HERE, "golden", "container_golden.json")))
    data = dict(container_inputs.determinism_inputs())
    L = refs.full("hip")
    was = K.set_cpu_fallback(True)
    try:
        for spec in ("launch:3", "sync:2:2", "copy:1", "sync:1:1000000000"):
            c0 = K.counters()
            for f in gold["frames"]:
                src = data[f["input"]]
                say("frames with", spec, f["input"], f["level"])
                K.fault_inject(spec)
                frame = refs.kolmo_compress(L, src, f["level"], threads=4)
                K.fault_inject(None)
                assert hashlib.sha256(frame.tobytes()).hexdigest() == f["sha256"], (spec, f)
                K.fault_inject(spec)
                back = refs.kolmo_decompress(L, frame, len(src), threads=2)
                K.fault_inject(None)
                assert back.tobytes() == src, (spec, f)
            c1 = K.counters()
            assert c1["cpu_fallbacks"] > c0["cpu_fallbacks"], spec
    finally:
        K.fault_inject(None)
        K.set_cpu_fallback(was)
    # and the device path afterwards, with nothing injected
    T = K.generate("mixed", 3, 1 << 20)
    r = 1 << K.codec18_rlog(T.size)
    U, I = K.bwt_aux(T, r)
    eU, eI = refs.oracle_bwt(T, r)
    assert np.array_equal(U, eU) and np.array_equal(I, eI)
