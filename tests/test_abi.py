"""The C-ABI library (CPU-side checks: no kernel runs here).

- libkbwt_hip.so loads and exports every function include/kbwt.h declares;
- argument validation answers exactly as kolmo's sais.c:12029-12037,12057-12068 does, before any device work;
- the n <= 1 cases of the contract, which need no device;
- a C caller compiled against include/kbwt.h links and runs the same validation (the drop-in seen from C).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import kolmo_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kbwt.h")).read()
    return re.findall(r"KBWT_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)


def test_exports_every_declared_symbol():
    names = declared_symbols()
    assert {"kolmo_sais_bwt_aux", "kolmo_sais_unbwt_aux", "kolmo_sais_unbwt"} <= set(names)
    L = K.lib()
    for name in names:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", K.HIP_LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported
    # nothing else leaks from the library's C++ internals
    assert all(not s.startswith("_Z") for s in exported)


def _call_bwt(T, U, A, n, r, I):
    L = K.lib()
    nil = ctypes.POINTER(ctypes.c_uint8)()
    niI = ctypes.POINTER(ctypes.c_int32)()
    p = lambda a, t=ctypes.POINTER(ctypes.c_uint8): a.ctypes.data_as(t) if a is not None else t()
    return L.kolmo_sais_bwt_aux(p(T), p(U), p(A, ctypes.POINTER(ctypes.c_int32)), n, r,
                                p(I, ctypes.POINTER(ctypes.c_int32)), 1)


def _call_unbwt(T, U, A, n, r, I):
    L = K.lib()
    p = lambda a, t=ctypes.POINTER(ctypes.c_uint8): a.ctypes.data_as(t) if a is not None else t()
    return L.kolmo_sais_unbwt_aux(p(T), p(U), p(A, ctypes.POINTER(ctypes.c_int32)), n, r,
                                  p(I, ctypes.POINTER(ctypes.c_int32)), 1)


def test_forward_validation():
    T = np.zeros(16, np.uint8); U = np.zeros(16, np.uint8); A = np.zeros(17, np.int32); I = np.zeros(16, np.int32)
    assert _call_bwt(None, U, A, 16, 2, I) == -1
    assert _call_bwt(T, None, A, 16, 2, I) == -1
    assert _call_bwt(T, U, None, 16, 2, I) == -1
    assert _call_bwt(T, U, A, 16, 2, None) == -1
    assert _call_bwt(T, U, A, -1, 2, I) == -1
    for r in (0, 1, 3, 6, -2):
        assert _call_bwt(T, U, A, 16, r, I) == -1


def test_forward_tiny_cases():
    T = np.array([7], np.uint8); U = np.zeros(1, np.uint8); A = np.zeros(2, np.int32); I = np.full(2, 99, np.int32)
    assert _call_bwt(T, U, A, 1, 2, I) == 0 and U[0] == 7 and I[0] == 1
    assert _call_bwt(T, U, A, 0, 2, I) == 0 and I[0] == 0


def test_inverse_validation():
    T = np.zeros(16, np.uint8); U = np.zeros(16, np.uint8); A = np.zeros(17, np.int32)
    I = np.ones(16, np.int32)
    assert _call_unbwt(None, U, A, 16, 4, I) == -1
    assert _call_unbwt(T, U, None, 16, 4, I) == -1
    assert _call_unbwt(T, U, A, 16, 3, I) == -1          # neither n nor a power of two
    assert _call_unbwt(T, U, A, 16, 1, I) == -1
    for bad in (0, -5, 17):
        J = I.copy(); J[2] = bad                           # I[t] outside 1..n for t <= (n-1)/r
        assert _call_unbwt(T, U, A, 16, 4, J) == -1
    J = I.copy(); J[4] = 0                                 # past (n-1)/r: never read
    assert _call_unbwt(T, U, A, 0, 4, J) == -1             # n = 0 needs I[0] == 0
    J[0] = 0
    assert _call_unbwt(T, U, A, 0, 4, J) == 0
    T1 = np.array([9], np.uint8); U1 = np.zeros(1, np.uint8)
    assert _call_unbwt(T1, U1, A, 1, 1, np.array([1], np.int32)) == 0 and U1[0] == 9
    assert _call_unbwt(T1, U1, A, 1, 1, np.array([2], np.int32)) == -1
    L = K.lib()
    p = lambda a, t=ctypes.POINTER(ctypes.c_uint8): a.ctypes.data_as(t)
    assert L.kolmo_sais_unbwt(p(T), p(U), p(A, ctypes.POINTER(ctypes.c_int32)), 16, 17, 1) == -1


def test_no_device_is_an_error_not_a_fallback():
    """Without a usable device the library answers -3; it never computes on the CPU."""
    if K.device_ok():
        pytest.skip("a device is present")
    T = np.frombuffer(b"abracadabra", np.uint8).copy()
    with pytest.raises(K.KbwtError) as e:
        K.bwt_aux(T, 2)
    assert e.value.code == K.KBWT_ERR_DEVICE


def test_c_caller_links(tmp_path):
    """A C program written against include/kbwt.h, as kolmo's bwt_stream.c would be, links with the library."""
    src = tmp_path / "caller.c"
    src.write_text(r'''
#include "kbwt.h"
#include <stdio.h>
int main(void) {
    uint8_t T[4] = {1, 2, 3, 4}, U[4]; int32_t A[5], I[4];
    if (kolmo_sais_bwt_aux(T, U, A, 4, 3, I, 1) != -1) return 1;          /* r not a power of two */
    if (kolmo_sais_bwt_aux(T, U, A, 1, 2, I, 1) != 0 || U[0] != 1 || I[0] != 1) return 2;
    I[0] = 5;
    if (kolmo_sais_unbwt(T, U, A, 4, 5, 1) != -1) return 3;               /* index out of range */
    printf("ok\n");
    return 0;
}
''')
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", K.LIBDIR, "-lkbwt_hip", f"-Wl,-rpath,{K.LIBDIR}"], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0 and res.stdout.strip() == "ok", res.stdout + res.stderr


def test_generators_deterministic():
    for kind in K.GEN_KINDS:
        a = K.generate(kind, 42, 100000)
        b = K.generate(kind, 42, 100000)
        c = K.generate(kind, 43, 100000)
        assert np.array_equal(a, b) and not np.array_equal(a, c), kind


def test_tuning_table_set_get_reset():
    """The per-size-class launch parameters (no device needed): the built-in table, a set, and a reset to it."""
    try:
        assert K.get_tuning(20) == (0, 3, 64)      # up to 1 MiB
        assert K.get_tuning(21) == (0, 4, 64)
        assert K.get_tuning(23) == (8, 4, 64)
        assert K.get_tuning(25) == (8, 6, 64)      # 32 MiB
        assert K.get_tuning(27) == (8, 7, 64)      # above 64 MiB
        K.set_tuning((5, 7, 64), 20)
        assert K.get_tuning(20) == (5, 7, 64) and K.get_tuning(21) == (0, 4, 64)
        K.reset_tuning(20)
        assert K.get_tuning(20) == (0, 3, 64)
        K.set_tuning((1, 9, 0))
        assert all(K.get_tuning(c) == (1, 9, 0) for c in range(33))
        with pytest.raises(K.KbwtError):
            K.set_tuning((11, 6, 16))              # no such radix configuration
        with pytest.raises(K.KbwtError):
            K.set_tuning((8, 13, 16))              # splitter spacing out of range
    finally:
        K.reset_tuning()
    assert K.get_tuning(22) == (8, 4, 64) and K.get_tuning(26) == (8, 6, 64) and K.get_tuning(30) == (8, 7, 64)


def test_concurrent_calls_without_device_fail_cleanly():
    """Concurrent kolmo_sais_* calls are gathered into batches by whichever caller finds none running. Without a
    device every batch fails, and every call (forward and inverse alike) must come back with -3, none left waiting."""
    if K.device_ok():
        pytest.skip("a device is present")
    import threading
    T = np.frombuffer(b"abracadabra" * 100, np.uint8).copy()
    codes, lock = [], threading.Lock()

    def call(k):
        for _ in range(5):
            try:
                if k % 2:
                    K.bwt_aux(T, 64)
                else:
                    K.unbwt_aux(T, 64, np.array([1] + [2] * 17, np.int32))
                rc = 0
            except K.KbwtError as e:
                rc = e.code
            with lock:
                codes.append(rc)

    ts = [threading.Thread(target=call, args=(k,)) for k in range(16)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in ts), "a caller is still waiting"
    assert len(codes) == 80 and set(codes) == {K.KBWT_ERR_DEVICE}


def test_block_workers_harness_loads_and_checks_arguments():
    """libkbwt_workers.so (kolmo's block workers as C threads) links against libkbwt_hip.so and rejects bad
    arguments before any call; without a device a pass reports its failed calls (-2) instead of a time."""
    W = K.workers_lib()
    T = np.frombuffer(b"abracadabra" * 100, np.uint8).copy()
    rs = np.array([64, 64], np.int32)
    run = W.kbwt_workers_run
    assert run(None, T.size, 600, K._ptr(rs, K._i32p), 2, 1) == -1.0
    assert run(K._ptr(T), T.size, 1, K._ptr(rs, K._i32p), 2, 1) == -1.0          # blocks of one byte
    assert run(K._ptr(T), T.size, 600, K._ptr(rs, K._i32p), 0, 1) == -1.0        # no workers
    assert run(K._ptr(T), 1099, 1098, K._ptr(rs, K._i32p), 2, 1) == -1.0         # a last block of one byte
    if not K.device_ok():
        assert run(K._ptr(T), T.size, 600, K._ptr(rs, K._i32p), 2, 1) == -2.0
        with pytest.raises(K.KbwtError):
            K.block_workers(T, 600, 2)
