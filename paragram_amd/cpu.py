"""The diagram on host cores: ctypes binding of libparagram_cpu.so
(include/paragram_cpu.h), the clipping algorithm of the HIP kernels in
C++/OpenMP.

``CpuBackend`` exposes the reference's two callables (the seam
paragram._backend._C, paragram/_backend.py:138-195; paragram/csrc/ext.cpp:10-29)
on CPU tensors, for machines without a GPU and as the CPU baseline of the
benchmark.  The product path (``paragram_amd.voronoi_diagram`` /
``power_diagram``) never calls it: it runs on the GPU or raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np
import torch

from . import _build

STAT_NAMES = ("fp64_rechecks", "exact", "unresolved", "clips", "candidates", "nodes", "build_ns", "index_ns",
              "cells_ns")


class CpuLibrary:
    def __init__(self, path: str = _build.CPU_LIB_PATH):
        if not os.path.exists(path):
            _build.build_cpu(lib_path=path)
        self.lib = lib = ctypes.CDLL(path)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        lib.pgc_abi_version.restype = ctypes.c_int
        lib.pgc_last_error.restype = ctypes.c_char_p
        lib.pgc_build.argtypes = [i64, vp, vp, i32, i32, ctypes.POINTER(vp)]
        lib.pgc_build.restype = ctypes.c_int
        lib.pgc_result_edges.argtypes = [vp]
        lib.pgc_result_edges.restype = i64
        lib.pgc_result_copy.argtypes = [vp, vp, vp, vp]
        lib.pgc_result_copy.restype = ctypes.c_int
        lib.pgc_result_stats.argtypes = [vp, vp]
        lib.pgc_result_stats.restype = ctypes.c_int
        lib.pgc_free.argtypes = [vp]
        lib.pgc_free.restype = None
        if lib.pgc_abi_version() != 1:
            raise ImportError("libparagram_cpu.so ABI version mismatch")

    def diagram(self, points, weights=None, leaf_size: int = -1, nthreads: int = 0):
        """``(adjacency, offsets, status, stats)`` as numpy int32 arrays (stats a dict)."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        n = pts.shape[0]
        h = ctypes.c_void_p()
        p = lambda a: None if a is None else ctypes.c_void_p(a.ctypes.data)
        if self.lib.pgc_build(n, p(pts), p(w), int(leaf_size), int(nthreads), ctypes.byref(h)) != 0:
            raise RuntimeError("pgc_build: " + self.lib.pgc_last_error().decode(errors="replace"))
        try:
            e = int(self.lib.pgc_result_edges(h))
            adj = np.zeros(max(e, 1), np.int32)
            off = np.zeros(n + 1, np.int32)
            st = np.zeros(max(n, 1), np.int32)
            if self.lib.pgc_result_copy(h, p(adj), p(off), p(st)) != 0:
                raise RuntimeError("pgc_result_copy: " + self.lib.pgc_last_error().decode(errors="replace"))
            stats = np.zeros(len(STAT_NAMES), np.int64)
            self.lib.pgc_result_stats(h, p(stats))
        finally:
            self.lib.pgc_free(h)
        return adj[:e], off, st[:n], dict(zip(STAT_NAMES, (int(x) for x in stats)))


class CpuBackend:
    """The reference's two callables on CPU tensors (same arguments and
    return values as paragram/csrc/ext.cpp:10-29)."""

    def __init__(self, lib: Optional[CpuLibrary] = None, nthreads: int = 0):
        self.lib = lib or CpuLibrary()
        self.nthreads = nthreads

    def _run(self, points, weights, knn, debug, initial_guesses, initial_guesses_offsets, leaf_size):
        # knn and initial guesses change the order of the clips, never the diagram
        del knn, debug, initial_guesses, initial_guesses_offsets
        p = points.detach().cpu().numpy()
        w = None if weights is None else weights.detach().cpu().numpy()
        adj, off, st, _ = self.lib.diagram(p, w, leaf_size, self.nthreads)
        return torch.from_numpy(adj), torch.from_numpy(off), torch.from_numpy(st)

    def build_voronoi_ultra(self, points, knn: int = 8, debug: bool = False, initial_guesses=None,
                            initial_guesses_offsets=None, leaf_size: int = -1):
        return self._run(points, None, knn, debug, initial_guesses, initial_guesses_offsets, leaf_size)

    def build_laguerre_voronoi_ultra(self, points, weights, knn: int = 8, debug: bool = False,
                                     initial_guesses=None, initial_guesses_offsets=None, leaf_size: int = -1):
        return self._run(points, weights, knn, debug, initial_guesses, initial_guesses_offsets, leaf_size)
