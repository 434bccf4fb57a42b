"""Loaders for the test oracle (oracle/liboracle.so) and for kolmo's own C compiled from source (oracle/_ref/).

Test infrastructure only: the product never loads these.
  oracle()       the C restatement: oracle_sa, oracle_bwt_aux, oracle_unbwt_aux, oracle_unbwt, rules
  ref()          kolmo's block-sorting path as built from the kolmo tree (sais.c and its codec glue, unchanged)
  codec(which)   kolmo's codec glue (bwt_stream.c, lzp.c, qlfc.c, par.c, klzma.c) with sais.c replaced:
                 which = "oracle" (over the oracle) or "hip" (over libkbwt_hip.so, the drop-in relink)
  full(which)    the whole kolmo library (kolmo_compress, kolmo_decompress): which = "ref" (unchanged), "oracle"
                 or "hip" (sais.c replaced as above)
"""
import ctypes
import hashlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REFDIR = os.path.join(ROOT, "oracle", "_ref")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_cache = {}


def _p(a, t=_u8p):
    return a.ctypes.data_as(t)


def _sais_sigs(L):
    L.kolmo_sais_bwt_aux.restype = ctypes.c_int32
    L.kolmo_sais_bwt_aux.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int32]
    L.kolmo_sais_unbwt_aux.restype = ctypes.c_int32
    L.kolmo_sais_unbwt_aux.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int32]
    L.kolmo_sais_unbwt.restype = ctypes.c_int32
    L.kolmo_sais_unbwt.argtypes = [_u8p, _u8p, _i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]


def _codec_sigs(L):
    L.kolmo_bsc_compress.restype = ctypes.c_int
    L.kolmo_bsc_compress.argtypes = [_u8p, _u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.kolmo_bsc_decompress.restype = ctypes.c_int
    L.kolmo_bsc_decompress.argtypes = [_u8p, ctypes.c_int, _u8p, ctypes.c_int, ctypes.c_int]
    L.kolmo_bwt_compress.restype = ctypes.c_int
    L.kolmo_bwt_compress.argtypes = [_u8p, _u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.kolmo_bwt_decompress.restype = ctypes.c_int
    L.kolmo_bwt_decompress.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_int, ctypes.c_int]
    return L


def oracle():
    if "oracle" not in _cache:
        L = ctypes.CDLL(ORACLE)
        L.oracle_sa.restype = ctypes.c_int32
        L.oracle_sa.argtypes = [_u8p, _i32p, ctypes.c_int32]
        L.oracle_bwt_aux.restype = ctypes.c_int32
        L.oracle_bwt_aux.argtypes = [_u8p, _u8p, ctypes.c_int32, ctypes.c_int32, _i32p]
        L.oracle_unbwt_aux.restype = ctypes.c_int32
        L.oracle_unbwt_aux.argtypes = [_u8p, _u8p, ctypes.c_int32, ctypes.c_int32, _i32p]
        L.oracle_unbwt.restype = ctypes.c_int32
        L.oracle_unbwt.argtypes = [_u8p, _u8p, ctypes.c_int32, ctypes.c_int32]
        L.oracle_bwt_mod.restype = ctypes.c_int32
        L.oracle_bwt_mod.argtypes = [ctypes.c_int32]
        L.oracle_codec18_rlog.restype = ctypes.c_int32
        L.oracle_codec18_rlog.argtypes = [ctypes.c_int32]
        _cache["oracle"] = L
    return _cache["oracle"]


def have_ref():
    return os.path.exists(os.path.join(REFDIR, "libkolmo_ref.so"))


def ref():
    if "ref" not in _cache:
        L = ctypes.CDLL(os.path.join(REFDIR, "libkolmo_ref.so"))
        _sais_sigs(L)
        _codec_sigs(L)
        _cache["ref"] = L
    return _cache["ref"]


def codec(which):
    key = "codec_" + which
    if key not in _cache:
        L = ctypes.CDLL(os.path.join(REFDIR, f"libkolmo_codec_{which}.so"))
        _sais_sigs(L)
        _codec_sigs(L)
        _cache[key] = L
    return _cache[key]


class KolmoParams(ctypes.Structure):
    """kolmo_params (src/kolmo.h:21-41)."""
    _fields_ = [("mode", ctypes.c_int), ("level", ctypes.c_int), ("threads", ctypes.c_int), ("block", ctypes.c_uint32),
                ("transforms", ctypes.c_int), ("mem_limit", ctypes.c_uint64)]


def have_full(which):
    return os.path.exists(os.path.join(REFDIR, f"libkolmo_full_{which}.so"))


def full(which):
    key = "full_" + which
    if key not in _cache:
        L = ctypes.CDLL(os.path.join(REFDIR, f"libkolmo_full_{which}.so"))
        L.kolmo_params_for_level.restype = ctypes.c_int
        L.kolmo_params_for_level.argtypes = [ctypes.c_int, ctypes.POINTER(KolmoParams)]
        L.kolmo_bound.restype = ctypes.c_size_t
        L.kolmo_bound.argtypes = [ctypes.c_size_t, ctypes.POINTER(KolmoParams)]
        L.kolmo_compress.restype = ctypes.c_int64
        L.kolmo_compress.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.POINTER(KolmoParams)]
        L.kolmo_decompress.restype = ctypes.c_int64
        L.kolmo_decompress.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.c_int]
        if which == "oracle":
            L.oracle_shim_calls.restype = ctypes.c_int64
            L.oracle_shim_calls.argtypes = [ctypes.c_int32]
        _cache[key] = L
    return _cache[key]


def kolmo_compress(L, src, level, threads=1, block=0):
    """A whole kolmo frame of src at a level (kolmo_params_for_level), with the given threads and block size."""
    src = as_u8(src)
    p = KolmoParams()
    assert L.kolmo_params_for_level(level, ctypes.byref(p)) == 0
    p.threads = threads
    p.block = block
    cap = L.kolmo_bound(src.size, ctypes.byref(p))
    out = np.empty(cap, np.uint8)
    k = L.kolmo_compress(_p(src), src.size, _p(out), cap, ctypes.byref(p))
    assert k > 0, k
    return out[:k]


def kolmo_decompress(L, frame, n, threads=1):
    frame = as_u8(frame)
    out = np.empty(max(n, 1), np.uint8)
    k = L.kolmo_decompress(_p(frame), frame.size, _p(out), out.size, threads)
    assert k == n, k
    return out[:n]


def as_u8(x):
    if isinstance(x, (bytes, bytearray)):
        return np.frombuffer(bytes(x), np.uint8).copy()
    return np.ascontiguousarray(x, np.uint8)


def nsamp(n, r):
    return (n - 1) // r + 1 if n > 0 else 1


# ---- oracle calls ----

def oracle_sa(T):
    T = as_u8(T)
    SA = np.empty(max(T.size, 1), np.int32)
    assert oracle().oracle_sa(_p(T), _p(SA, _i32p), T.size) == 0
    return SA[: T.size]


def oracle_bwt(T, r):
    T = as_u8(T)
    n = T.size
    U = np.empty(max(n, 1), np.uint8)
    I = np.zeros(nsamp(n, r), np.int32)
    rc = oracle().oracle_bwt_aux(_p(T), _p(U), n, r, _p(I, _i32p))
    assert rc == 0, rc
    return U[:n], I


def oracle_unbwt(U, r, I):
    U = as_u8(U)
    n = U.size
    T = np.empty(max(n, 1), np.uint8)
    I = np.ascontiguousarray(I, np.int32)
    rc = oracle().oracle_unbwt_aux(_p(U), _p(T), n, r, _p(I, _i32p))
    assert rc == 0, rc
    return T[:n]


# ---- calls into a library with kolmo's sais.h entry points (the reference, or a relinked codec) ----

def sais_bwt(L, T, r, threads=1):
    T = as_u8(T)
    n = T.size
    U = np.empty(max(n, 1), np.uint8)
    A = np.empty(max(n, 1) + 1, np.int32)
    I = np.zeros(nsamp(n, r), np.int32)
    rc = L.kolmo_sais_bwt_aux(_p(T), _p(U), _p(A, _i32p), n, r, _p(I, _i32p), threads)
    return rc, U[:n], I


def sais_unbwt(L, U, r, I, threads=1):
    U = as_u8(U)
    n = U.size
    T = np.empty(max(n, 1), np.uint8)
    A = np.empty(max(n, 1) + 1, np.int32)
    I = np.ascontiguousarray(I, np.int32)
    rc = L.kolmo_sais_unbwt_aux(_p(U), _p(T), _p(A, _i32p), n, r, _p(I, _i32p), threads)
    return rc, T[:n]


def bsc_compress(L, T, lzp_hash=15, lzp_min=72, coder=1, threads=1):
    """kolmo_bsc_compress (src/codecs/bwt_stream.h): coder 1 = static QLFC, 2 = adaptive."""
    T = as_u8(T)
    out = np.empty(T.size + 64, np.uint8)
    rc = L.kolmo_bsc_compress(_p(T), _p(out), T.size, lzp_hash, lzp_min, coder, threads)
    assert rc > 0, rc
    return out[:rc]


class BwtParams(ctypes.Structure):
    _fields_ = [("lzp_hash", ctypes.c_int), ("lzp_min", ctypes.c_int), ("coder", ctypes.c_int),
                ("index_log2", ctypes.c_int), ("min_block", ctypes.c_int), ("static_pct", ctypes.c_int)]


BWT_DEFAULTS = (17, 96, 2, 0, 4 << 20, 0)   # KOLMO_BWT_DEFAULTS, src/codecs/bwt_stream.h:37


def bwt18_compress(L, T, params=BWT_DEFAULTS, threads=1):
    T = as_u8(T)
    cap = T.size + 64 + 4 * 256
    out = np.empty(cap, np.uint8)
    p = BwtParams(*params)
    rc = L.kolmo_bwt_compress(_p(T), _p(out), cap, T.size, ctypes.byref(p), threads)
    return rc, out[: max(rc, 0)]


def bwt18_decompress(L, payload, n, threads=1):
    payload = as_u8(payload)
    out = np.empty(n + 1, np.uint8)
    rc = L.kolmo_bwt_decompress(_p(payload), payload.size, _p(out), n, threads)
    return rc, out[:n]


def fnv1a(b):
    h = 1469598103934665603
    for x in bytes(b):
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
