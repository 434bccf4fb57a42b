"""Host-side C++ of the library built for the CPU alone and run under sanitizers (no device needed).

kolmo runs its own tests under ASan/UBSan (Makefile:162); the host parts of this library that involve threads are
built here from their own headers with g++ and -fsanitize=thread or address,undefined, and stressed:

- kbwt_copypool.h: the thread pool that fills and drains the pinned staging chunks (consecutive lists of unequal
  part counts from several callers, every byte checked);
- kbwt_combine.h: the queue that gathers concurrent kolmo_sais_* calls into batches, driven by a fake batch runner
  (every caller gets its own block's result, failures stay per block, oversized blocks are refused, nobody waits
  forever).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "kolmo_amd", "csrc")
GXX = shutil.which("g++")

pytestmark = pytest.mark.skipif(GXX is None, reason="g++ not available")


def build_and_run(src, out, san, args=(), timeout=240):
    exe = os.path.join(out, os.path.basename(src).replace(".cpp", f"_{san.replace(',', '_')}"))
    subprocess.run([GXX, "-O1", "-g", "-std=c++17", f"-fsanitize={san}", "-fno-omit-frame-pointer", "-pthread",
                    "-I", CSRC, "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1 print_stacktrace=1")
    res = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout, env=env)
    assert res.returncode == 0, f"{os.path.basename(exe)} rc {res.returncode}\n{res.stdout[-3000:]}{res.stderr[-6000:]}"
    return res.stdout


def test_copypool_tsan(tmp_path):
    out = build_and_run(os.path.join(HERE, "native", "copypool_stress.cpp"), str(tmp_path), "thread", [25])
    assert "copypool ok" in out


def test_copypool_asan(tmp_path):
    out = build_and_run(os.path.join(HERE, "native", "copypool_stress.cpp"), str(tmp_path), "address,undefined", [25])
    assert "copypool ok" in out


def test_combine_tsan(tmp_path):
    out = build_and_run(os.path.join(HERE, "native", "combine_stress.cpp"), str(tmp_path), "thread", [80])
    assert "combine ok" in out


def test_combine_asan(tmp_path):
    out = build_and_run(os.path.join(HERE, "native", "combine_stress.cpp"), str(tmp_path), "address,undefined", [80])
    assert "combine ok" in out
