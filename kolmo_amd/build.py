"""Build recipe for the package's native libraries (in-tree, under kolmo_amd/lib/).

- libkbwt_hip.so: the HIP kernels and the C-ABI of include/kbwt.h, compiled for gfx950 by hipcc.
- libkbwt_gen.so: the seeded input generators (host C).
- libkbwt_workers.so: kolmo's block workers as C threads over libkbwt_hip.so, for measuring concurrent calls.

`python -m kolmo_amd.build` runs it; __graft_entry__.build() calls build_all().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["kbwt_kernels.hip", "kbwt_rounds.hip", "kbwt_router.hip", "kbwt_host.hip", "kbwt_router_host.hip",
               "kbwt_cpu.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-fvisibility=hidden",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:2]))
    return res


def build_hip(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libkbwt_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs, verbose)
    os.replace(tmp, out)
    return out


def build_gen(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libkbwt_gen.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-fopenmp", "-o", tmp, os.path.join(CSRC, "kbwt_gen.c"), "-lm",
          "-lpthread"], verbose)
    os.replace(tmp, out)
    return out


def build_workers(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libkbwt_workers.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-fvisibility=hidden", "-I", os.path.join(ROOT, "include"),
          "-o", tmp, os.path.join(CSRC, "kbwt_workers.c"), "-L", LIB, "-lkbwt_hip", "-Wl,-rpath,$ORIGIN",
          "-lpthread"], verbose)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    return [build_hip(verbose), build_gen(verbose), build_workers(verbose)]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
