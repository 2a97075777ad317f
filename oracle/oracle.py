"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by the test suite, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s CPU-baseline leg.  The product package never imports it.
Each function names the reference code it restates (see ``oracle.h``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        u32 = ctypes.c_uint32
        L.o_orient3d_sign_batch.argtypes = [P, i64, P]
        L.o_insphere_sign_batch.argtypes = [P, i64, P]
        L.o_orient3d_sos_batch.argtypes = [P, P, i64, P, P]
        L.o_orient4d_batch.argtypes = [P, P, i64, P, P]
        L.o_orient2d_plain.argtypes = [P, P, P]
        L.o_orient2d_plain.restype = ctypes.c_double
        L.o_orient3d_lifted_sign.argtypes = [P, P, P, P]
        L.o_orient2d_lifted_sign.argtypes = [P, P, P]
        L.o_orient1d_lifted_sign.argtypes = [P, P]
        L.o_normalize_points.argtypes = [P, i64, P, P, P]
        L.o_morton_code.argtypes = [P, P, P]
        L.o_morton_code.restype = u32
        L.o_expand_bits.argtypes = [u32]
        L.o_expand_bits.restype = u32
        L.o_initial_stride.argtypes = [i64]
        L.o_initial_stride.restype = i64
        L.o_stratified_order.argtypes = [i64, i64, P]
        L.o_delaunay_bw.argtypes = [P, i64, ctypes.POINTER(ctypes.POINTER(u32)), P]
        L.o_delaunay_bw.restype = i64
        L.o_delaunay_bw_reference.argtypes = [P, i64, ctypes.POINTER(ctypes.POINTER(u32))]
        L.o_delaunay_bw_reference.restype = i64
        L.o_delaunay_bw_many.argtypes = [P, P, i64, ctypes.c_int, P]
        L.o_delaunay_bw_many.restype = i64
        L.o_free.argtypes = [P]
        L.o_canonical_real.argtypes = [P, i64, u32, P]
        L.o_canonical_real.restype = i64
        L.o_rebuild_adjacency.argtypes = [P, i64, P]
        L.o_count_violations.argtypes = [P, i64, P, P, i64, ctypes.c_int]
        L.o_count_violations.restype = i64
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# -- exact predicates (src/predicates.rs:14-21) ------------------------------

def orient3d_sign(q: np.ndarray) -> np.ndarray:
    """q: (k, 4, 3) -> sign of det[a-d; b-d; c-d] per row."""
    q = _f64(q).reshape(-1, 12)
    out = np.empty(len(q), np.int32)
    lib().o_orient3d_sign_batch(_p(q), len(q), _p(out))
    return out


def insphere_sign(q: np.ndarray) -> np.ndarray:
    """q: (k, 5, 3) -> sign of Shewchuk's insphere per row."""
    q = _f64(q).reshape(-1, 15)
    out = np.empty(len(q), np.int32)
    lib().o_insphere_sign_batch(_p(q), len(q), _p(out))
    return out


def orient3d_sos(q: np.ndarray, idx: np.ndarray, with_depth: bool = False):
    """src/predicates.rs:106-208; +1 when the perturbed raw det is < 0."""
    q = _f64(q).reshape(-1, 12)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 4)
    out = np.empty(len(q), np.int32)
    dep = np.empty(len(q), np.int32)
    lib().o_orient3d_sos_batch(_p(q), _p(idx), len(q), _p(out), _p(dep))
    return (out, dep) if with_depth else out


def orient4d(q: np.ndarray, idx: np.ndarray, with_depth: bool = False):
    """src/predicates.rs:450-675; +1 outside / -1 inside (raw-negative tets)."""
    q = _f64(q).reshape(-1, 15)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 5)
    out = np.empty(len(q), np.int32)
    dep = np.empty(len(q), np.int32)
    lib().o_orient4d_batch(_p(q), _p(idx), len(q), _p(out), _p(dep))
    return (out, dep) if with_depth else out


def orient2d_plain(a, b, c) -> float:
    a, b, c = (_f64(x) for x in (a, b, c))
    return lib().o_orient2d_plain(_p(a), _p(b), _p(c))


def orient3d_lifted_sign(a, b, c, d) -> int:
    a, b, c, d = (_f64(x) for x in (a, b, c, d))
    return lib().o_orient3d_lifted_sign(_p(a), _p(b), _p(c), _p(d))


def orient2d_lifted_sign(a, b, c) -> int:
    a, b, c = (_f64(x) for x in (a, b, c))
    return lib().o_orient2d_lifted_sign(_p(a), _p(b), _p(c))


def orient1d_lifted_sign(a, b) -> int:
    a, b = (_f64(x) for x in (a, b))
    return lib().o_orient1d_lifted_sign(_p(a), _p(b))


# -- point-set helpers --------------------------------------------------------

def normalize_points(xyz: np.ndarray):
    """src/lib.rs:313-337 -> (normalized f32 (n,3), min (3,), scale)."""
    x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.empty_like(x)
    mn = np.empty(3, np.float32)
    sc = np.empty(1, np.float32)
    lib().o_normalize_points(_p(x), len(x), _p(out), _p(mn), _p(sc))
    return out, mn, float(sc[0])


def morton_code(p, mn, mx) -> int:
    p, mn, mx = (np.ascontiguousarray(v, dtype=np.float32) for v in (p, mn, mx))
    return int(lib().o_morton_code(_p(p), _p(mn), _p(mx)))


def expand_bits(v: int) -> int:
    return int(lib().o_expand_bits(v))


def initial_stride(n: int) -> int:
    return int(lib().o_initial_stride(n))


def stratified_order(n: int, stride: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().o_stratified_order(n, stride, _p(out))
    return out


# -- triangulation -------------------------------------------------------------

def delaunay(points_normalized: np.ndarray, return_stats: bool = False):
    """SoS Delaunay of the normalized points + 4 super vertices (ids n..n+3).

    Returns all tets (m, 4) uint32, raw orient3d <= 0 (gDel3D orientation).
    """
    pts = np.ascontiguousarray(points_normalized, dtype=np.float32).reshape(-1, 3)
    ptr = ctypes.POINTER(ctypes.c_uint32)()
    stats = np.zeros(4, np.int64)
    m = lib().o_delaunay_bw(_p(pts), len(pts), ctypes.byref(ptr), _p(stats))
    if m < 0:
        raise RuntimeError("oracle Bowyer-Watson failed (inconsistent cavity)")
    tets = np.ctypeslib.as_array(ptr, shape=(m * 4,)).copy().reshape(m, 4)
    lib().o_free(ctypes.cast(ptr, ctypes.c_void_p))
    return (tets, stats) if return_stats else tets


def delaunay_reference(points_normalized: np.ndarray) -> np.ndarray:
    """full_bw_reconstruction as written (src/phase2.rs:383-606): input
    order, linear seed scan, append-only tets.  O(N^2); same contract as
    ``delaunay``."""
    pts = np.ascontiguousarray(points_normalized, dtype=np.float32).reshape(-1, 3)
    ptr = ctypes.POINTER(ctypes.c_uint32)()
    m = lib().o_delaunay_bw_reference(_p(pts), len(pts), ctypes.byref(ptr))
    if m < 0:
        raise RuntimeError("reference Bowyer-Watson restatement failed")
    tets = np.ctypeslib.as_array(ptr, shape=(m * 4,)).copy().reshape(m, 4)
    lib().o_free(ctypes.cast(ptr, ctypes.c_void_p))
    return tets


def delaunay_many(points_normalized: np.ndarray, offsets, threads: int) -> np.ndarray:
    """``delaunay`` of each cloud [offsets[c], offsets[c+1]) of already
    per-cloud-normalized points, the clouds spread over `threads` OpenMP
    threads; returns the tet count of each cloud."""
    pts = np.ascontiguousarray(points_normalized, dtype=np.float32).reshape(-1, 3)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    counts = np.zeros(len(off) - 1, np.int64)
    if lib().o_delaunay_bw_many(_p(pts), _p(off), len(off) - 1, int(threads), _p(counts)) < 0:
        raise RuntimeError("oracle Bowyer-Watson failed on a cloud")
    return counts


def canonical_real(tets: np.ndarray, n: int) -> np.ndarray:
    """Real-only tets (all ids < n), each quad sorted, list sorted."""
    t = np.ascontiguousarray(tets, dtype=np.uint32).reshape(-1, 4)
    out = np.empty_like(t)
    k = lib().o_canonical_real(_p(t), len(t), n, _p(out))
    return out[:k].copy()


def rebuild_adjacency(tets: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(tets, dtype=np.uint32).reshape(-1, 4)
    adj = np.empty_like(t)
    lib().o_rebuild_adjacency(_p(t), len(t), _p(adj))
    return adj


def count_violations(points_normalized, tets, adj, only_real: bool = True) -> int:
    pts = np.ascontiguousarray(points_normalized, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tets, dtype=np.uint32).reshape(-1, 4)
    a = np.ascontiguousarray(adj, dtype=np.uint32).reshape(-1, 4)
    return int(lib().o_count_violations(_p(pts), len(pts), _p(t), _p(a), len(t),
                                        1 if only_real else 0))


def extended_points(points_normalized: np.ndarray) -> np.ndarray:
    """Normalized points + super vertices + infinity as f64 (n+5, 3)."""
    p = np.asarray(points_normalized, dtype=np.float32).astype(np.float64)
    sup = np.array([[-100, -100, -100], [400, -100, -100], [-100, 400, -100],
                    [-100, -100, 400], [1000, 1000, 1000]], np.float64)
    return np.vstack([p, sup])
