"""Test tool: the reference's own FP32 cell construction restated on the CPU
(refsim.cpp; see its header for what is restated and what stands in), and
the count of the cells where its output departs from the exact diagram, by
reason.  Built with g++ into tests/refsim/librefsim.so on first use."""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librefsim.so")
EVICTED, TOO_FAR_BEHIND, DROPPED_ZERO_DIFF, PLANE_FULL, VERTEX_FULL, WALK_FAILED = 1, 2, 4, 8, 16, 32
_lib = None


def build() -> str:
    cxx = os.environ.get("CXX") or shutil.which("g++")
    cmd = [cxx, "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           os.path.join(HERE, "refsim.cpp"), "-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("refsim build failed:\n" + r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        src = os.path.join(HERE, "refsim.cpp")
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
            build()
        _lib = ctypes.CDLL(LIB)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        _lib.refsim_run.argtypes = [i64, vp, vp, i32, i32, vp, i64, vp, vp, vp, vp, i64, i32]
        _lib.refsim_run.restype = i64
    return _lib


def run(points, weights=None, cells=None, knn: int = 8, leaf: int = 10, nthreads: int = 0):
    """Raw rows of the restated reference: (counts, status, flags, offsets, flat)."""
    lib = _load()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(-1)
    n = pts.shape[0]
    sel = None if cells is None else np.ascontiguousarray(cells, np.int64)
    m = n if sel is None else sel.shape[0]
    counts = np.zeros(m, np.int32)
    status = np.zeros(m, np.int32)
    flags = np.zeros(m, np.int32)
    cap = 64 * max(m, 1)
    flat = np.zeros(cap, np.int32)
    p = lambda a: None if a is None else ctypes.c_void_p(a.ctypes.data)
    e = lib.refsim_run(n, p(pts), p(w), knn, leaf, p(sel), m, p(counts), p(status), p(flags), p(flat), cap, nthreads)
    if e < 0:
        raise RuntimeError("refsim: row capacity exceeded")
    off = np.zeros(m + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    return counts, status, flags, off, flat[:e]


def divergence(points, weights=None, cells=None, nthreads: int = 0) -> dict:
    """Cells (raw rows and statuses) where the restated reference differs from
    the oracle, by first cause: the 64-plane or 64-vertex capacity (status 1
    or 2 written at some point; after one, later boundary walks on the
    broken cell can end in status 5, and the last status written is the one
    reported), the boundary walk's step limit (status 5), a hidden cell
    missed or wrongly reported (status 3 on one side only) with a skipped
    point behind the seed (the unsigned too-far test) or without one, stack
    eviction, and other row differences (FP32 predicates)."""
    import oracle

    counts, status, flags, off, flat = run(points, weights, cells, nthreads=nthreads)
    ref = oracle.cells(points, weights, cells, nthreads=nthreads)
    reasons = {"status_1_plane_capacity": 0, "status_2_vertex_capacity": 0, "status_5_boundary_walk": 0,
               "hidden_mismatch_too_far_behind": 0, "hidden_mismatch_other": 0, "stack_eviction": 0,
               "rows_fp32_predicates": 0}
    differ = 0
    evicted = int(((flags & EVICTED) != 0).sum())
    for k in range(counts.shape[0]):
        row = flat[off[k]:off[k + 1]]
        if status[k] == ref.status[k] and np.array_equal(row, ref.row(k)):
            continue
        differ += 1
        if flags[k] & PLANE_FULL:
            reasons["status_1_plane_capacity"] += 1
        elif flags[k] & VERTEX_FULL:
            reasons["status_2_vertex_capacity"] += 1
        elif flags[k] & WALK_FAILED:
            reasons["status_5_boundary_walk"] += 1
        elif (status[k] == 3) != (ref.status[k] == 3):
            key = "hidden_mismatch_too_far_behind" if flags[k] & TOO_FAR_BEHIND else "hidden_mismatch_other"
            reasons[key] += 1
        elif flags[k] & EVICTED:
            reasons["stack_eviction"] += 1
        else:
            reasons["rows_fp32_predicates"] += 1
    return {"checked": int(counts.shape[0]), "differ": differ, "reasons": reasons,
            "cells_with_eviction": evicted, "status_counts": {int(s): int((status == s).sum()) for s in np.unique(status)}}
