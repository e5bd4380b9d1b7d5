"""The whole pipeline of one call captured in a HIP graph (single cloud).

``GraphedDiagram(n)`` allocates every buffer of a call for n points once,
captures index build -> kNN pre-clip -> cells -> offsets -> symmetry check ->
CSR emission into one graph, and ``__call__(points[, weights])`` replays it
for new coordinates of the same shape: one graph launch and one
device-to-host read (the edge count) per diagram, no allocation.  The rare
union pass (an asymmetric edge) runs eagerly after the replay.

No counterpart in the reference (its calls are eager); the
result is that of ``voronoi_diagram`` / ``power_diagram`` on the same input.
Buffers are cleared by the library's own fill kernel inside the graph (no
memset nodes).
"""
from __future__ import annotations

import ctypes

import torch

from . import _engine

_p = _engine._p


class GraphedDiagram:
    def __init__(self, n: int, weighted: bool = False, knn: int = 8, leaf_size: int = -1,
                 device: torch.device | str = "cuda"):
        from ._backend import _C, _load_error

        if _C is None:
            raise ImportError(f"paragram_amd HIP engine unavailable: {_load_error}")
        if n < 1:
            raise ValueError("a graphed diagram needs at least one point")
        lib = self.lib = _C.lib
        dev = self.dev = torch.device(device)
        self.n = n
        i32 = dict(dtype=torch.int32, device=dev)
        self.points = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.weights = torch.zeros((n, 1), dtype=torch.float32, device=dev) if weighted else None
        ws_bytes = self.ws_bytes = lib.u64("pg_index_workspace_bytes", n)
        self.ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        waves = ctypes.c_int32(0)
        lib.call("pg_default_waves", ctypes.byref(waves))
        self.waves = int(waves.value)
        self.scratch_bytes = lib.u64("pg_cell_scratch_bytes", n)
        self.scratch = torch.empty(self.scratch_bytes, dtype=torch.uint8, device=dev)
        cap = lib.u64("pg_row_pool_elems", n)
        self.rows = torch.empty(cap, **i32)
        self.row_offsets = torch.empty(n, dtype=torch.int64, device=dev)
        self.counts = torch.empty(n, **i32)
        self.status = torch.empty(n, **i32)
        self.stats = torch.empty(_engine.NUM_STATS, dtype=torch.int64, device=dev)
        self.offsets = torch.empty(n + 1, **i32)
        self.adjacency = torch.empty(max(cap, 1), **i32)  # every row fits: E <= the row pool
        self.scan_tmp = torch.empty(max(lib.u64("pg_scan_tmp_bytes", n + 1), 4), dtype=torch.uint8, device=dev)
        self.stride = _engine.KNN_PRECLIP if knn > 0 else 0
        self.guesses = torch.empty(max(n * self.stride, 1), **i32)
        self.guess_d2 = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
        self.leaf_size = leaf_size
        self._capture()

    def _capture(self):
        self._pipeline()  # once eagerly (first-use work outside the capture)
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._pipeline()
        torch.cuda.synchronize(self.dev)

    def _pipeline(self):
        lib, n = self.lib, self.n
        sp = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        lib.call("pg_fill_u32", _p(self.stats), 2 * _engine.NUM_STATS, 0, sp)
        lib.call("pg_build_index", n, _p(self.points), _p(self.weights), int(self.leaf_size), _p(self.ws),
                 self.ws_bytes, sp)
        g = gd = None
        if self.stride:
            lib.call("pg_knn_guesses", n, _p(self.ws), 0, self.stride, _p(self.guesses), _p(self.guess_d2), sp)
            g, gd = self.guesses, self.guess_d2
        lib.call("pg_compute_cells_ex", n, _p(self.ws), 0, n, _p(g), None, 0, self.stride, _p(gd), 0, None, None,
                 _p(self.rows), self.rows.numel(), _p(self.row_offsets), _p(self.counts), _p(self.status), _p(self.stats),
                 _p(self.scratch), self.scratch_bytes, self.waves, None, sp)
        lib.call("pg_scan_counts", n, _p(self.counts), _p(self.offsets), _p(self.scan_tmp), sp)
        lib.call("pg_check_symmetry", n, _p(self.ws), _p(self.counts), _p(self.rows), _p(self.row_offsets), 0,
                 _p(self.offsets), _p(self.stats), sp)
        lib.call("pg_emit_csr", n, _p(self.ws), _p(self.counts), _p(self.offsets), _p(self.rows),
                 _p(self.row_offsets), 0, _p(self.adjacency), sp)

    def __call__(self, points: torch.Tensor, weights: torch.Tensor | None = None):
        """``(adjacency, offsets, status)`` of the new points (views into this object's buffers,
        valid until the next call)."""
        if tuple(points.shape) != (self.n, 3):
            raise ValueError(f"points must have shape ({self.n}, 3)")
        if (weights is None) != (self.weights is None):
            raise ValueError("weights must be given exactly when the graph was built for weights")
        self.points.copy_(points)
        if weights is not None:
            self.weights.copy_(weights.reshape(self.n, 1))
        self.graph.replay()
        st = self.stats.cpu()  # the one device-to-host read
        if int(st[_engine.STAT_POOL_SHORT]):
            # the rows did not fit the row pool: grow it to what the launch
            # reported, capture again and replay (the only allocating path)
            need = int(st[_engine.STAT_POOL_DEMAND])
            i32 = dict(dtype=torch.int32, device=self.dev)
            self.rows = self.adjacency = None
            self.rows = torch.empty(need + need // 8 + 65536, **i32)
            self.adjacency = torch.empty(self.rows.numel(), **i32)
            self._capture()
            self.graph.replay()
            st = self.stats.cpu()
        e = int(st[15])
        adjacency, offsets = self.adjacency[:e], self.offsets
        if int(st[8]) or int(st[9]):
            sp = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
            adjacency, offsets = _engine._union(self.lib, self.n, self.counts, offsets, adjacency, self.scan_tmp,
                                                sp, self.dev)
        return adjacency, offsets, self.status


class GraphedBatch:
    """A batch of clouds of fixed sizes captured in one HIP graph: batch index
    build -> kNN pre-clip -> cells -> offsets -> symmetry check -> CSR
    emission.  ``__call__(points[, weights])`` copies the new coordinates in,
    replays, and makes one device-to-host read (the statistics with E and
    every cloud's entry count); it returns one ``(adjacency, offsets, status)``
    per cloud, equal to ``batched_voronoi_diagram`` / ``batched_power_diagram``
    of the same input (cloud-local ids; fresh tensors of a few small kernels
    after the replay).

    No counterpart in the reference (one cloud per call there).  The cloud
    offsets are checked on the host and written to the workspace once
    (``pg_batch_set_offsets``); inside the graph the library launches kernels
    only (no copies, no memset nodes).
    """

    def __init__(self, cloud_offsets, weighted: bool = False, knn: int = 8, leaf_size: int = -1,
                 device: torch.device | str = "cuda"):
        import numpy as np

        from ._api import _validate_cloud_offsets
        from ._backend import _C, _load_error

        if _C is None:
            raise ImportError(f"paragram_amd HIP engine unavailable: {_load_error}")
        off = np.asarray(cloud_offsets)
        n = int(off[-1]) if off.size else 0
        self.clouds = _validate_cloud_offsets(off, n)
        if n < 1:
            raise ValueError("a graphed batch needs at least one point")
        lib = self.lib = _C.lib
        dev = self.dev = torch.device(device)
        self.n, self.n_clouds = n, int(self.clouds.shape[0]) - 1
        i32 = dict(dtype=torch.int32, device=dev)
        self.points = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.weights = torch.zeros((n, 1), dtype=torch.float32, device=dev) if weighted else None
        ws_bytes = self.ws_bytes = lib.u64("pg_batch_workspace_bytes", n, self.n_clouds)
        self.ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        with torch.cuda.device(dev):
            sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
            lib.call("pg_batch_set_offsets", n, self.n_clouds, ctypes.c_void_p(self.clouds.ctypes.data), _p(self.ws),
                     ws_bytes, sp)
        waves = ctypes.c_int32(0)
        lib.call("pg_default_waves", ctypes.byref(waves))
        self.waves = int(waves.value)
        self.scratch_bytes = lib.u64("pg_cell_scratch_bytes", n)
        self.scratch = torch.empty(self.scratch_bytes, dtype=torch.uint8, device=dev)
        self.rows = torch.empty(lib.u64("pg_row_pool_elems", n), **i32)
        self.adjacency = torch.empty(self.rows.numel(), **i32)
        self.row_offsets = torch.empty(n, dtype=torch.int64, device=dev)
        self.counts = torch.empty(n, **i32)
        self.status = torch.empty(n, **i32)
        # statistics then the CSR offset at each cloud start: one read per call
        self.readback = torch.empty(_engine.NUM_STATS + self.n_clouds + 1, dtype=torch.int64, device=dev)
        self.stats = self.readback[:_engine.NUM_STATS]
        self.cloud_starts = torch.from_numpy(self.clouds.copy()).to(dev)
        self.offsets = torch.empty(n + 1, **i32)
        self.scan_tmp = torch.empty(max(lib.u64("pg_scan_tmp_bytes", n + 1), 4), dtype=torch.uint8, device=dev)
        self.stride = _engine.KNN_PRECLIP if knn > 0 else 0
        self.guesses = torch.empty(max(n * self.stride, 1), **i32)
        self.guess_d2 = torch.empty(n, dtype=torch.float32, device=dev)
        self.leaf_size = leaf_size
        self._capture()

    def _capture(self):
        self._pipeline()
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._pipeline()
        torch.cuda.synchronize(self.dev)

    def _pipeline(self):
        lib, n, nc = self.lib, self.n, self.n_clouds
        sp = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        lib.call("pg_fill_u32", _p(self.stats), 2 * _engine.NUM_STATS, 0, sp)
        lib.call("pg_build_batch_index_resident", n, nc, _p(self.points), _p(self.weights), int(self.leaf_size),
                 _p(self.ws), self.ws_bytes, sp)
        g = gd = None
        if self.stride:
            lib.call("pg_knn_guesses", n, _p(self.ws), nc, self.stride, _p(self.guesses), _p(self.guess_d2), sp)
            g, gd = self.guesses, self.guess_d2
        lib.call("pg_compute_cells_batch_resident", n, _p(self.ws), nc, _p(g), None, self.stride, _p(gd),
                 _p(self.rows), self.rows.numel(), _p(self.row_offsets), _p(self.counts), _p(self.status),
                 _p(self.stats), _p(self.scratch), self.scratch_bytes, self.waves, None, sp)
        lib.call("pg_scan_counts", n, _p(self.counts), _p(self.offsets), _p(self.scan_tmp), sp)
        lib.call("pg_check_symmetry", n, _p(self.ws), _p(self.counts), _p(self.rows), _p(self.row_offsets), 0,
                 _p(self.offsets), _p(self.stats), sp)
        lib.call("pg_emit_csr", n, _p(self.ws), _p(self.counts), _p(self.offsets), _p(self.rows),
                 _p(self.row_offsets), 0, _p(self.adjacency), sp)
        self.readback[_engine.NUM_STATS:].copy_(self.offsets.index_select(0, self.cloud_starts))

    def __call__(self, points: torch.Tensor, weights: torch.Tensor | None = None):
        """One ``(adjacency, offsets, status)`` per cloud."""
        if tuple(points.shape) != (self.n, 3):
            raise ValueError(f"points must have shape ({self.n}, 3)")
        if (weights is None) != (self.weights is None):
            raise ValueError("weights must be given exactly when the graph was built for weights")
        self.points.copy_(points)
        if weights is not None:
            self.weights.copy_(weights.reshape(self.n, 1))
        self.graph.replay()
        rb = self.readback.cpu()  # the one device-to-host read
        st = rb[:_engine.NUM_STATS]
        if int(st[_engine.STAT_POOL_SHORT]):
            # the rows did not fit the row pool: grow it to what the launch
            # reported, capture again and replay (the only allocating path)
            need = int(st[_engine.STAT_POOL_DEMAND])
            i32 = dict(dtype=torch.int32, device=self.dev)
            self.rows = self.adjacency = None
            self.rows = torch.empty(need + need // 8 + 65536, **i32)
            self.adjacency = torch.empty(self.rows.numel(), **i32)
            self._capture()
            self.graph.replay()
            rb = self.readback.cpu()
            st = rb[:_engine.NUM_STATS]
        e = int(st[15])
        adjacency, offsets = self.adjacency[:e], self.offsets
        at = rb[_engine.NUM_STATS:].numpy()
        entries = at[1:] - at[:-1]
        if int(st[8]) or int(st[9]):
            sp = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
            adjacency, offsets = _engine._union(self.lib, self.n, self.counts, offsets, adjacency, self.scan_tmp,
                                                sp, self.dev)
            entries = None  # the union changed the rows: split_batch reads the counts
        return _engine.split_batch(adjacency, offsets, self.status, self.clouds, entries)
