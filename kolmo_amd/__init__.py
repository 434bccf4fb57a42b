"""kolmo_amd: kolmo's block-sorting path (forward BWT with sampled indexes, inverse BWT) on MI355X.

The product is the C-ABI library kolmo_amd/lib/libkbwt_hip.so (include/kbwt.h). This module is its host-side mirror
for Python callers and tests. The raw entry points keep the names, argument meaning and return codes of
src/algo/sais/sais.h:13,17,20 of kolmo 0.0.2 (kolmo_sais_bwt_aux, kolmo_sais_unbwt_aux, kolmo_sais_unbwt); the
convenience wrappers bwt_aux / unbwt_aux raise KbwtError on a non-zero code.

There is no CPU path in this module: loading fails loudly when the library is missing, and a call fails with code -3
when no usable device is present. The library's own CPU path after a device error is opt-in
(set_cpu_fallback(True), or KBWT_CPU_FALLBACK=1) and every use of it is counted (counters()).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIBDIR, "libkbwt_hip.so")
GEN_LIB = os.path.join(LIBDIR, "libkbwt_gen.so")
WORKERS_LIB = os.path.join(LIBDIR, "libkbwt_workers.so")

KBWT_OK, KBWT_ERR_ARG, KBWT_ERR_NOMEM, KBWT_ERR_DEVICE = 0, -1, -2, -3
MAX_ROUNDS = 40


class KbwtError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"kbwt error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("blocks", ctypes.c_int32), ("sigma", ctypes.c_int32),
        ("bits_per_symbol", ctypes.c_int32),
        ("h0", ctypes.c_int32), ("rounds", ctypes.c_int32), ("passes", ctypes.c_int32 * MAX_ROUNDS),
        ("active", ctypes.c_int64 * MAX_ROUNDS), ("splitters", ctypes.c_int32),
        ("ms_fwd", ctypes.c_double), ("ms_inv", ctypes.c_double),
    ]

    def as_dict(self):
        r = self.rounds
        return {"n": self.n, "blocks": self.blocks, "sigma": self.sigma, "bits_per_symbol": self.bits_per_symbol, "h0": self.h0,
                "rounds": r, "passes": list(self.passes[: r + 1]), "active": list(self.active[: r + 1]),
                "splitters": self.splitters, "ms_fwd": self.ms_fwd, "ms_inv": self.ms_inv}


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("ms", ctypes.c_double), ("launches", ctypes.c_int64),
                ("elems", ctypes.c_int64)]


class Counters(ctypes.Structure):
    """kbwt_counters: process-wide counts of the error path and of combined batches."""
    _fields_ = [("device_errors", ctypes.c_int64), ("retried_blocks", ctypes.c_int64),
                ("cpu_fallbacks", ctypes.c_int64), ("injected_faults", ctypes.c_int64),
                ("combined_batches", ctypes.c_int64), ("combined_blocks", ctypes.c_int64),
                ("graph_replays", ctypes.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Seg(ctypes.Structure):
    """kbwt_seg: one block of a batch (byte offset in the batch buffers, length, sampling interval)."""
    _fields_ = [("offset", ctypes.c_int64), ("n", ctypes.c_int32), ("r", ctypes.c_int32)]


_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None
_gen = None


def lib():
    """The loaded libkbwt_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise ImportError(f"{HIP_LIB} is missing: build it with `python -m kolmo_amd.build`")
        # One HIP runtime per process: torch wheels bundle their own libamdhip64 (same soname), and a second copy
        # loaded next to ROCm's breaks device initialisation of whichever comes second. Loading torch first lets this
        # library bind to the runtime torch already holds.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = ctypes.CDLL(HIP_LIB)
        for name in ("kolmo_sais_bwt_aux", "kolmo_sais_unbwt_aux"):
            f = getattr(L, name)
            f.restype = ctypes.c_int32
        L.kolmo_sais_bwt_aux.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int32]
        L.kolmo_sais_unbwt_aux.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int32]
        L.kolmo_sais_unbwt.restype = ctypes.c_int32
        L.kolmo_sais_unbwt.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.kbwt_dev_bwt_aux.restype = ctypes.c_int32
        L.kbwt_dev_bwt_aux.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32p,
                                       ctypes.c_void_p]
        L.kbwt_dev_unbwt_aux.restype = ctypes.c_int32
        L.kbwt_dev_unbwt_aux.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32p,
                                         ctypes.c_void_p]
        L.kbwt_dev_unbwt_aux_noalloc.restype = ctypes.c_int32
        L.kbwt_dev_unbwt_aux_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                                 _i32p, ctypes.c_void_p]
        L.kbwt_last_stats.restype = ctypes.c_int32
        L.kbwt_last_stats.argtypes = [ctypes.POINTER(Stats)]
        L.kbwt_last_error.restype = ctypes.c_char_p
        L.kbwt_device_ok.restype = ctypes.c_int32
        L.kbwt_release.restype = None
        L.kbwt_profile.restype = None
        L.kbwt_profile.argtypes = [ctypes.c_int32]
        L.kbwt_profile_only.restype = None
        L.kbwt_profile_only.argtypes = [ctypes.c_char_p]
        L.kbwt_profile_read.restype = ctypes.c_int32
        L.kbwt_profile_read.argtypes = [ctypes.POINTER(KernelTime), ctypes.c_int32]
        for name in ("kbwt_bwt_aux_batch", "kbwt_unbwt_aux_batch"):
            f = getattr(L, name)
            f.restype = ctypes.c_int32
            f.argtypes = [_u8p, _u8p, ctypes.POINTER(Seg), ctypes.c_int32, _i32p]
        for name in ("kbwt_dev_bwt_aux_batch", "kbwt_dev_unbwt_aux_batch"):
            f = getattr(L, name)
            f.restype = ctypes.c_int32
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Seg), ctypes.c_int32, _i32p,
                          ctypes.c_void_p]
        L.kbwt_set_cpu_fallback.restype = ctypes.c_int32
        L.kbwt_set_cpu_fallback.argtypes = [ctypes.c_int32]
        L.kbwt_fault_inject.restype = ctypes.c_int32
        L.kbwt_fault_inject.argtypes = [ctypes.c_char_p]
        L.kbwt_get_counters.restype = ctypes.c_int32
        L.kbwt_get_counters.argtypes = [ctypes.POINTER(Counters)]
        L.kbwt_dev_copy.restype = ctypes.c_int32
        L.kbwt_dev_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        _lib = L
    return _lib


_workers = None


def workers_lib():
    """libkbwt_workers.so: kolmo's block workers as C threads over libkbwt_hip.so (loaded first)."""
    global _workers
    if _workers is None:
        if not os.path.exists(WORKERS_LIB):
            raise ImportError(f"{WORKERS_LIB} is missing: build it with `python -m kolmo_amd.build`")
        lib()
        W = ctypes.CDLL(WORKERS_LIB)
        W.kbwt_workers_run.restype = ctypes.c_double
        W.kbwt_workers_run.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int32, _i32p, ctypes.c_int32, ctypes.c_int32]
        W.kbwt_workers_last_error.restype = ctypes.c_char_p
        W.kbwt_workers_last_error.argtypes = []
        _workers = W
    return _workers


def block_workers(T, block, workers, reps=1):
    """Best seconds over `reps` passes of `workers` C threads over T cut into `block`-byte blocks, each thread taking
    blocks in turn and calling kolmo_sais_bwt_aux then kolmo_sais_unbwt_aux (host buffers, codec 18's r per block),
    every round trip checked against its block. Raises KbwtError when a call fails or a round trip differs."""
    import numpy as np
    T = np.ascontiguousarray(T, dtype=np.uint8)
    sizes = [min(block, T.size - o) for o in range(0, T.size, block)]
    rs = np.array([1 << codec18_rlog(m) for m in sizes], dtype=np.int32)
    sec = workers_lib().kbwt_workers_run(_ptr(T), T.size, block, _ptr(rs, _i32p), workers, reps)
    if sec < 0:
        raise KbwtError(int(sec), "block workers: bad arguments" if sec == -1 else
                        "block workers: " + workers_lib().kbwt_workers_last_error().decode())
    return sec


def gen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_LIB):
            raise ImportError(f"{GEN_LIB} is missing: build it with `python -m kolmo_amd.build`")
        _gen = ctypes.CDLL(GEN_LIB)
    return _gen


def _ptr(a, t=_u8p):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise KbwtError(rc, lib().kbwt_last_error().decode())


def n_samples(n, r):
    return (n - 1) // r + 1 if n > 0 else 1


def bwt_aux(T, r):
    """BWT of T (bytes or uint8 array) with sampling interval r: returns (U as uint8 array, I as int32 array)."""
    T = np.ascontiguousarray(np.frombuffer(T, np.uint8) if isinstance(T, (bytes, bytearray)) else T, np.uint8)
    n = T.size
    U = np.empty(max(n, 1), np.uint8)
    I = np.zeros(n_samples(n, r), np.int32)
    A = np.empty(1, np.int32)   # caller scratch of the reference contract; unused by the device path
    _check(lib().kolmo_sais_bwt_aux(_ptr(T), _ptr(U), _ptr(A, _i32p), n, r, _ptr(I, _i32p), 1))
    return U[:n], I


def unbwt_aux(U, r, I):
    """Inverse BWT: returns the original bytes as a uint8 array."""
    U = np.ascontiguousarray(np.frombuffer(U, np.uint8) if isinstance(U, (bytes, bytearray)) else U, np.uint8)
    I = np.ascontiguousarray(I, np.int32)
    n = U.size
    T = np.empty(max(n, 1), np.uint8)
    A = np.empty(1, np.int32)
    _check(lib().kolmo_sais_unbwt_aux(_ptr(U), _ptr(T), _ptr(A, _i32p), n, r, _ptr(I, _i32p), 1))
    return T[:n]


def dev_bwt_aux(dT, r, dU=None, stream=None):
    """Device-resident forward on a uint8 torch tensor of the current device: returns (dU, I)."""
    import torch
    n = dT.numel()
    if dU is None:
        dU = torch.empty_like(dT)
    I = np.zeros(n_samples(n, r), np.int32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().kbwt_dev_bwt_aux(dT.data_ptr(), dU.data_ptr(), n, r, _ptr(I, _i32p), s))
    return dU, I


def dev_unbwt_aux(dU, r, I, dT=None, stream=None):
    import torch
    n = dU.numel()
    if dT is None:
        dT = torch.empty_like(dU)
    I = np.ascontiguousarray(I, np.int32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().kbwt_dev_unbwt_aux(dU.data_ptr(), dT.data_ptr(), n, r, _ptr(I, _i32p), s))
    return dT


def dev_unbwt_aux_noalloc(dU, r, I, dT=None, stream=None):
    """kbwt_dev_unbwt_aux_noalloc: the device-resident inverse from the library's static scratch (n <= 4 MiB)."""
    import torch
    n = dU.numel()
    if dT is None:
        dT = torch.empty_like(dU)
    I = np.ascontiguousarray(I, np.int32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().kbwt_dev_unbwt_aux_noalloc(dU.data_ptr(), dT.data_ptr(), n, r, _ptr(I, _i32p), s))
    return dT


def make_segs(sizes, rs, gaps=None):
    """Segments for blocks of the given sizes packed back to back (optionally with gap bytes after each)."""
    segs = (Seg * max(len(sizes), 1))()
    at = 0
    for b, (n, r) in enumerate(zip(sizes, rs)):
        segs[b].offset, segs[b].n, segs[b].r = at, int(n), int(r)
        at += int(n) + (int(gaps[b]) if gaps is not None else 0)
    return segs, at


def samples_total(sizes, rs):
    return sum(n_samples(int(n), int(r)) for n, r in zip(sizes, rs))


def bwt_aux_batch(blocks, rs):
    """Forward BWT of a list of blocks (uint8 arrays) in one batch: returns (list of U, list of I)."""
    sizes = [int(b.size) for b in blocks]
    segs, total = make_segs(sizes, rs)
    T = np.concatenate([np.asarray(b, np.uint8) for b in blocks]) if blocks else np.empty(0, np.uint8)
    T = np.ascontiguousarray(np.concatenate([T, np.zeros(1, np.uint8)]))
    U = np.empty_like(T)
    I = np.zeros(max(samples_total(sizes, rs), 1), np.int32)
    _check(lib().kbwt_bwt_aux_batch(_ptr(T), _ptr(U), segs, len(blocks), _ptr(I, _i32p)))
    outU, outI, k = [], [], 0
    for b, n in enumerate(sizes):
        outU.append(U[segs[b].offset: segs[b].offset + n].copy())
        ns = n_samples(n, rs[b])
        outI.append(I[k: k + ns].copy())
        k += ns
    return outU, outI


def unbwt_aux_batch(blocks, rs, Is):
    sizes = [int(b.size) for b in blocks]
    segs, total = make_segs(sizes, rs)
    T = np.concatenate([np.asarray(b, np.uint8) for b in blocks]) if blocks else np.empty(0, np.uint8)
    T = np.ascontiguousarray(np.concatenate([T, np.zeros(1, np.uint8)]))
    U = np.empty_like(T)
    I = np.ascontiguousarray(np.concatenate(list(Is)) if Is else np.zeros(1, np.int32), np.int32)
    _check(lib().kbwt_unbwt_aux_batch(_ptr(T), _ptr(U), segs, len(blocks), _ptr(I, _i32p)))
    return [U[segs[b].offset: segs[b].offset + n].copy() for b, n in enumerate(sizes)]


def dev_bwt_aux_batch(dT, segs, count, dU=None, stream=None):
    """Device-resident batch forward over a uint8 tensor holding the blocks at segs' offsets: returns (dU, I)."""
    import torch
    if dU is None:
        dU = torch.empty_like(dT)
    I = np.zeros(max(sum(n_samples(segs[b].n, segs[b].r) for b in range(count)), 1), np.int32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().kbwt_dev_bwt_aux_batch(dT.data_ptr(), dU.data_ptr(), segs, count, _ptr(I, _i32p), s))
    return dU, I


def dev_unbwt_aux_batch(dU, segs, count, I, dT=None, stream=None):
    import torch
    if dT is None:
        dT = torch.empty_like(dU)
    I = np.ascontiguousarray(I, np.int32)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().kbwt_dev_unbwt_aux_batch(dU.data_ptr(), dT.data_ptr(), segs, count, _ptr(I, _i32p), s))
    return dT


def last_stats():
    st = Stats()
    lib().kbwt_last_stats(ctypes.byref(st))
    return st.as_dict()


def device_ok():
    return bool(lib().kbwt_device_ok())


def counters():
    """kbwt_get_counters: device errors, blocks retried alone, CPU-path fallbacks, injected faults, combined batches."""
    c = Counters()
    _check(lib().kbwt_get_counters(ctypes.byref(c)))
    return c.as_dict()


def set_cpu_fallback(on):
    """kbwt_set_cpu_fallback: the library's CPU path for host-pointer calls whose device work fails. Returns the
    previous setting."""
    return bool(lib().kbwt_set_cpu_fallback(1 if on else 0))


def fault_inject(spec):
    """kbwt_fault_inject: 'site:k[:count]' (site alloc, launch, sync or copy) fails those occurrences of the fault
    point as if the HIP runtime had; None or '' disarms."""
    rc = lib().kbwt_fault_inject(spec.encode() if spec else None)
    if rc != 0:
        raise KbwtError(rc, f"bad fault spec {spec!r}")


# ---- sampling-interval rules of the callers (src/codecs/bwt_stream.c:40-44 and :224-228) ----

def bwt_mod(n):
    """codec 5: the largest power of two <= n/8, minus one (0 when n < 16)."""
    x = n // 8
    if x <= 0:
        return 0
    return (1 << (x.bit_length() - 1)) - 1


def codec18_rlog(n):
    """codec 18 with index_log2 = 0: log2 of the sampling interval, at most 255 sampled indexes."""
    r = bwt_mod(n) + 1
    rlog = r.bit_length() - 1
    rlog = max(rlog, 1)
    while rlog < 30 and ((n - 1) >> rlog) > 255:
        rlog += 1
    return rlog


# ---- seeded inputs ----

GEN_KINDS = {"markov2": 0, "mixed": 1, "repeats": 2, "dna": 3, "random": 4, "csvcols": 5, "logs": 6, "bf16": 7,
             "text": 8}


def generate(kind, seed, n, model_seed=None):
    """Seeded synthetic block (see kolmo_amd/csrc/kbwt_gen.c for each generator). model_seed (default: seed) draws
    the Markov models, so blocks of one batch can share their source statistics."""
    g = gen_lib()
    g.kbwt_gen_kind2.restype = ctypes.c_int
    g.kbwt_gen_kind2.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _u8p, ctypes.c_size_t]
    out = np.empty(max(n, 1), np.uint8)
    rc = g.kbwt_gen_kind2(GEN_KINDS[kind] if isinstance(kind, str) else int(kind),
                          seed if model_seed is None else model_seed, seed, _ptr(out), n)
    if rc != 0:
        raise RuntimeError(f"generator {kind} failed ({rc})")
    return out[:n]


def generate_bsc_kind(kind, n):
    """The input kinds of kolmo's tests/test_bwt.c:24-45."""
    g = gen_lib()
    g.kbwt_gen_bsc_kind.restype = ctypes.c_int
    g.kbwt_gen_bsc_kind.argtypes = [ctypes.c_int, _u8p, ctypes.c_int]
    out = np.empty(max(n, 1), np.uint8)
    g.kbwt_gen_bsc_kind(kind, _ptr(out), n)
    return out[:n]


def batch_plan(seed, count, lo, hi, kinds):
    g = gen_lib()
    g.kbwt_gen_batch_plan.restype = ctypes.c_int
    kinds = np.asarray([GEN_KINDS[k] if isinstance(k, str) else int(k) for k in kinds], np.int32)
    sizes = np.empty(count, np.int64)
    kout = np.empty(count, np.int32)
    seeds = np.empty(count, np.uint64)
    rc = g.kbwt_gen_batch_plan(ctypes.c_uint64(seed), ctypes.c_int(count), ctypes.c_int64(lo), ctypes.c_int64(hi),
                               kinds.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(kinds)),
                               sizes.ctypes.data_as(ctypes.c_void_p), kout.ctypes.data_as(ctypes.c_void_p),
                               seeds.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError("bad batch plan")
    return sizes, kout, seeds


def generate_batch(seed, count, lo, hi, kinds, threads=8):
    """A batch of seeded blocks: sizes log-uniform in [lo, hi], each from one of `kinds` (Markov models drawn from
    the batch seed). Returns (list of uint8 arrays, kinds, seeds)."""
    from concurrent.futures import ThreadPoolExecutor
    sizes, kind_of, seeds = batch_plan(seed, count, lo, hi, kinds)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        blocks = list(ex.map(lambda b: generate(int(kind_of[b]), int(seeds[b]), int(sizes[b]), model_seed=seed),
                             range(count)))
    return blocks, kind_of, seeds


# The measured workloads: (generator, seed, n).
WORKLOADS = {
    "cfg1_markov2_1MiB": ("markov2", 1, 1 << 20),
    "cfg2_mixed_32MiB": ("mixed", 2, 32 << 20),
    "cfg3_repeats_128MiB": ("repeats", 3, 128 << 20),
}


# ---- autotuner (kbwt_autotune): launch parameters per size class, none of which changes any output ----

class Tuning(ctypes.Structure):
    _fields_ = [("sort_cfg", ctypes.c_int32), ("inv_dshift", ctypes.c_int32), ("seg_rank", ctypes.c_int32),
                ("ms", ctypes.c_float)]




def _tune_sigs(L):
    L.kbwt_autotune.restype = ctypes.c_int32
    L.kbwt_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(Seg), ctypes.c_int32, ctypes.c_void_p,
                                ctypes.POINTER(Tuning), ctypes.POINTER(Tuning)]
    L.kbwt_set_tuning.restype = ctypes.c_int32
    L.kbwt_set_tuning.argtypes = [ctypes.c_int32, ctypes.POINTER(Tuning)]
    L.kbwt_get_tuning.restype = ctypes.c_int32
    L.kbwt_get_tuning.argtypes = [ctypes.c_int32, ctypes.POINTER(Tuning)]
    return L


def autotune(dT, segs, count, start=None, stream=None):
    """Time the candidate launch parameters on these device blocks and install the fastest for their size class.
    Returns (best (sort_cfg, inv_dshift, seg_rank), best ms, ms at the start point)."""
    import torch
    L = _tune_sigs(lib())
    best = Tuning()
    st = Tuning(*start, 0.0) if start is not None else Tuning(-1, 0, 0, 0.0)   # not a valid setting: the built-in entry
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(L.kbwt_autotune(dT.data_ptr(), segs, count, s, ctypes.byref(best), ctypes.byref(st)))
    return (best.sort_cfg, best.inv_dshift, best.seg_rank), best.ms, st.ms


def set_tuning(t, size_log2=-1):
    _check(_tune_sigs(lib()).kbwt_set_tuning(size_log2, ctypes.byref(Tuning(*t, 0.0))))


def reset_tuning(size_log2=-1):
    """Back to the library's built-in table (one size class, or all)."""
    _check(_tune_sigs(lib()).kbwt_set_tuning(size_log2, None))


def get_tuning(size_log2):
    t = Tuning()
    _check(_tune_sigs(lib()).kbwt_get_tuning(size_log2, ctypes.byref(t)))
    return (t.sort_cfg, t.inv_dshift, t.seg_rank)
