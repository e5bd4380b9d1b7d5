"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the CPU oracle (oracle.cpp).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker.  The product
(``paragram_amd``) never imports it.

The oracle decides every clipping predicate exactly (FP64 filter, Shewchuk
expansion fallback) and returns, per cell, the sorted neighbour row and the
status code of the reference (0 success, 3 empty_cell;
voronoi_convex_cell.cuh:91-100).  ``diagram`` adds the union symmetrisation of
adjacency.cuh:78-159 and returns the same (adjacency, offsets, status) triple
as the reference's ``build_voronoi_ultra`` (voronoi_ultra.cu:143-376).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

STAT_NAMES = ("fp64_decided", "exact_fallbacks", "exact_ties", "duplicates", "zero_orientations",
              "clips", "nodes", "candidates")


def build() -> str:
    """Compile oracle.cpp with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64, i32, vp = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p
    lib.oracle_run.argtypes = [i64, vp, vp, ctypes.c_int, vp, i64]
    lib.oracle_run.restype = vp
    lib.oracle_result_size.argtypes = [vp]
    lib.oracle_result_size.restype = i64
    lib.oracle_result_copy.argtypes = [vp, vp, vp, vp, vp]
    lib.oracle_result_copy.restype = None
    lib.oracle_free.argtypes = [vp]
    lib.oracle_free.restype = None
    lib.oracle_symmetrize.argtypes = [i64, vp, vp, vp, vp, vp, i64]
    lib.oracle_symmetrize.restype = i64
    lib.oracle_exact_side.argtypes = [i64, vp, vp, i64, vp, i32, vp, vp]
    lib.oracle_exact_side.restype = ctypes.c_int
    lib.oracle_box.argtypes = [i64, vp, vp, vp]
    lib.oracle_box.restype = None
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _prep(points, weights):
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        assert w.shape[0] == pts.shape[0]
    return pts, w


class Rows:
    """Raw rows of some cells: CSR ``(counts, offsets int64, flat)`` plus status."""

    def __init__(self, counts, status, flat, stats):
        self.counts = counts
        self.status = status
        self.flat = flat
        self.offsets = np.zeros(counts.shape[0] + 1, np.int64)
        np.cumsum(counts, out=self.offsets[1:])
        self.stats = stats

    def __len__(self):
        return self.counts.shape[0]

    def row(self, k):
        return self.flat[self.offsets[k]:self.offsets[k + 1]]


def cells(points, weights=None, cells=None, nthreads: int = 0) -> Rows:
    """Raw (unsymmetrised) sorted rows and statuses of ``cells`` (all cells when None)."""
    lib = _load()
    pts, w = _prep(points, weights)
    n = pts.shape[0]
    sel = None if cells is None else np.ascontiguousarray(cells, dtype=np.int64)
    m = n if sel is None else sel.shape[0]
    h = lib.oracle_run(n, _ptr(pts), _ptr(w), nthreads, _ptr(sel), m)
    try:
        e = lib.oracle_result_size(h)
        counts = np.zeros(m, np.int32)
        status = np.zeros(m, np.int32)
        flat = np.zeros(max(e, 1), np.int32)
        stats = np.zeros(8, np.int64)
        lib.oracle_result_copy(h, _ptr(counts), _ptr(status), _ptr(flat), _ptr(stats))
    finally:
        lib.oracle_free(h)
    return Rows(counts, status, flat[:e], dict(zip(STAT_NAMES, (int(x) for x in stats))))


def symmetrize(rows: Rows):
    """Union symmetrisation with empty-cell drop; returns (adjacency, offsets)."""
    lib = _load()
    n = len(rows)
    offsets = np.zeros(n + 1, np.int32)
    adj_cap = 2 * int(rows.counts.sum()) + 1
    adj = np.zeros(adj_cap, np.int32)
    flat = rows.flat if rows.flat.size else np.zeros(1, np.int32)
    e = lib.oracle_symmetrize(n, _ptr(rows.counts), _ptr(rows.offsets), _ptr(flat), _ptr(offsets), _ptr(adj),
                              adj_cap)
    if e < 0:
        raise RuntimeError("oracle: adjacency capacity exceeded")
    return adj[:e].copy(), offsets


def diagram(points, weights=None, nthreads: int = 0):
    """Full oracle diagram: ``(adjacency, offsets, status, stats)`` as numpy int32."""
    r = cells(points, weights, None, nthreads)
    if len(r) == 0:
        return np.zeros(0, np.int32), np.zeros(1, np.int32), r.status, r.stats
    adj, offsets = symmetrize(r)
    return adj, offsets, r.status, r.stats


def exact_side(points, weights, seed: int, tri, q: int):
    """Exact sign(P)*sign(D) for vertex ``tri`` of the seed's cell against plane ``q``.

    Plane ids >= 0 are point ids; -1-(2*axis+side) are box planes (side 0 = low).
    Returns ``(side, sign_D, sign_P)``.
    """
    lib = _load()
    pts, w = _prep(points, weights)
    t = np.ascontiguousarray(tri, dtype=np.int32)
    sd = ctypes.c_int32(0)
    sp = ctypes.c_int32(0)
    r = lib.oracle_exact_side(pts.shape[0], _ptr(pts), _ptr(w), seed, _ptr(t), q, ctypes.byref(sd), ctypes.byref(sp))
    return int(r), int(sd.value), int(sp.value)


def box(points):
    """FP32 box (lo, hi) exactly as the clipper uses it."""
    lib = _load()
    pts, _ = _prep(points, None)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib.oracle_box(pts.shape[0], _ptr(pts), _ptr(lo), _ptr(hi))
    return lo, hi
