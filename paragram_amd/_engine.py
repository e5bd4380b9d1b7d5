"""Host driver: runs the HIP stages of libparagram_hip.so on torch's stream.

This is the MI355X counterpart of the reference's C++ orchestration
``build_voronoi_ultra`` (paragram/csrc/laguerre/voronoi_ultra.cu:143-376) and
``build_laguerre_voronoi_ultra`` (laguerre_ultra.cu:176-454): index build,
chunked cell construction, CSR assembly.  All device memory comes from the
torch caching allocator; all work is ordered on
``torch.cuda.current_stream()`` of the points' device; the only
device-to-host read is one 192-byte counter block that carries E (and the
row-pool counters: in the rare case that the rows did not fit, the cells run
again with the pool size the launch reported).
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np
import torch

from ._chunking import plan_chunks

NUM_STATS = 32
STAT_NAMES = {
    0: "fp64_rechecks", 1: "unresolved", 2: "inconsistent", 3: "promoted", 4: "clips",
    5: "candidates", 6: "nodes", 7: "cells", 8: "asymmetric", 9: "dropped", 10: "leaves",
    11: "skipped", 12: "truncated", 13: "pool_short", 14: "pool_demand", 15: "edges", 16: "exact",
    24: "bail_test", 25: "bail_vertex", 26: "tentative", 27: "tentative_wrong",
}
STAT_POOL_SHORT, STAT_POOL_DEMAND = 13, 14
STAT_EXACT_LOG = 17  # 17..23: original ids + 1 of the first cells with an expansion-arithmetic test
STAT_TENTATIVE_LOG = 28  # 28..31: original ids + 1 of the first cells rebuilt after a wrong FP32 sign
CELL_BLOCK = 64  # cells per admission block (one wavefront)
KNN_PRECLIP = 16  # nearest neighbours clipped first when knn > 0

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparagram_hip.so")


class HipLibrary:
    """ctypes binding of include/paragram_hip.h."""

    def __init__(self, path: str = LIB_PATH):
        self.path = path
        self.lib = lib = ctypes.CDLL(path)
        i64, i32, u64p, vp = ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p
        u64 = ctypes.c_uint64
        sig = {
            "pg_abi_version": ([], ctypes.c_int),
            "pg_last_error": ([], ctypes.c_char_p),
            "pg_index_workspace_bytes": ([i64, u64p], ctypes.c_int),
            "pg_build_index": ([i64, vp, vp, i32, vp, ctypes.c_uint64, vp], ctypes.c_int),
            "pg_default_waves": ([ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
            "pg_cell_scratch_bytes": ([i64, u64p], ctypes.c_int),
            "pg_row_pool_elems": ([i64, u64p], ctypes.c_int),
            "pg_compute_cells": ([i64, vp, i64, i64, vp, vp, i32, i32, vp, u64, vp, vp, vp, vp, vp, u64, i32, vp, vp],
                                 ctypes.c_int),
            "pg_scan_tmp_bytes": ([i64, u64p], ctypes.c_int),
            "pg_exact_side": ([i64, vp, i64, vp, vp, vp], ctypes.c_int),
            "pg_exact_side_host": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_exact_vertex_host": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_batch_localize": ([i64, i32, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_index_layout": ([i64, u64p], ctypes.c_int),
            "pg_batch_set_offsets": ([i64, i32, vp, vp, u64, vp], ctypes.c_int),
            "pg_build_batch_index_resident": ([i64, i32, vp, vp, i32, vp, u64, vp], ctypes.c_int),
            "pg_compute_cells_batch_resident": ([i64, vp, i32, vp, vp, i32, vp, vp, u64, vp, vp, vp, vp, vp, u64,
                                                 i32, vp, vp], ctypes.c_int),
            "pg_leaf_stats": ([i64, vp, vp, vp], ctypes.c_int),
            "pg_noalloc_max_cells": ([ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
            "pg_compute_cells_noalloc": ([i64, vp, i64, i64, vp, vp, i32, i32, vp, i32, vp, vp, vp, u64, vp, vp, vp,
                                          vp, i32, vp, vp], ctypes.c_int),
            "pg_scan_counts": ([i64, vp, vp, vp, vp], ctypes.c_int),
            "pg_chunk_counts": ([i64, vp, i64, i64, vp, vp, vp], ctypes.c_int),
            "pg_pack_rows": ([i64, vp, i64, i64, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_check_symmetry": ([i64, vp, vp, vp, vp, i64, vp, vp, vp], ctypes.c_int),
            "pg_emit_csr": ([i64, vp, vp, vp, vp, vp, i64, vp, vp], ctypes.c_int),
            "pg_symmetrize_union": ([i64, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp], ctypes.c_int),
            "pg_fill_u32": ([vp, i64, ctypes.c_uint32, vp], ctypes.c_int),
            "pg_knn": ([i64, vp, i32, vp, vp, vp], ctypes.c_int),
            "pg_knn_batch": ([i64, vp, i32, i32, vp, vp, vp], ctypes.c_int),
            "pg_knn_guesses": ([i64, vp, i32, i32, vp, vp, vp], ctypes.c_int),
            "pg_mlp_param_count": ([], ctypes.c_int),
            "pg_mlp_features": ([i64, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_mlp_features_batch": ([i64, vp, i32, vp, vp, vp, vp, vp], ctypes.c_int),
            "pg_mlp_forward": ([i64, vp, vp, vp, vp], ctypes.c_int),
            "pg_mlp_train_scratch_elems": ([i64, u64p], ctypes.c_int),
            "pg_mlp_train_step": ([i64, vp, vp, vp, ctypes.c_float, vp, vp, vp, vp, vp, ctypes.c_float, i32, vp],
                                  ctypes.c_int),
            "pg_compute_cells_ex": ([i64, vp, i64, i64, vp, vp, i32, i32, vp, i32, vp, vp, vp, u64, vp, vp, vp, vp, vp,
                                     u64, i32, vp, vp], ctypes.c_int),
            "pg_batch_workspace_bytes": ([i64, i32, u64p], ctypes.c_int),
            "pg_build_batch_index": ([i64, i32, vp, vp, vp, i32, vp, ctypes.c_uint64, vp], ctypes.c_int),
            "pg_batch_quads": ([i32, vp, i64, i64, vp], ctypes.c_int),
            "pg_compute_cells_batch": ([i64, vp, i32, vp, i64, i64, vp, vp, i32, vp, vp, u64, vp, vp, vp, vp, vp,
                                        u64, i32, vp, vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if lib.pg_abi_version() != 2:
            raise ImportError("libparagram_hip.so ABI version mismatch")

    def call(self, name: str, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            msg = self.lib.pg_last_error().decode(errors="replace")
            raise RuntimeError(msg if msg.startswith("Kernel failed") else f"{name}: {msg}")

    def u64(self, name: str, *args) -> int:
        out = ctypes.c_uint64(0)
        self.call(name, *args, ctypes.byref(out))
        return int(out.value)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Instrument:
    """Optional per-call measurement hooks (events on the launch stream)."""

    def __init__(self):
        self.cell_events = []  # (start, end) torch.cuda.Event pairs around each cell launch
        self.stats = None
        self.plan = None
        self.cycles = None     # set to an int32 tensor [n] to collect per-cell clock cycles


def build_diagram(lib: HipLibrary, points: torch.Tensor, weights, knn: int = 8, debug: bool = False,
                  initial_guesses=None, initial_guesses_offsets=None, leaf_size: int = -1,
                  instrument: Instrument | None = None, clouds=None, init_h=None, extent=None,
                  radius_model=None, preclip: int | None = None, row_pool: int | None = None,
                  waves_per_cu: int | None = None):
    """Returns ``(adjacency int32[E], offsets int32[N+1], status int32[N])`` on points' device.

    ``clouds``: optional host int64 array of cloud offsets (a batch of clouds,
    contiguous ranges of ``points``); the result is then the batch as one
    diagram over all points, neighbour ids global.
    ``init_h``: optional float32 [N] predicted half-width of each cell (the
    output does not depend on it); ``extent``: optional float32 [N] that
    receives the built half-width of each cell.  ``radius_model``: a
    ``learned.RadiusModel`` whose prediction (kNN, features and network on
    this call's index) supplies ``init_h``.  ``preclip``: nearest neighbours
    clipped first when knn > 0 (default KNN_PRECLIP, 1..16).  ``row_pool``:
    initial row-pool elements per launch (default ``pg_row_pool_elems``);
    a pool that proves too small grows, so only the time depends on it.
    ``waves_per_cu``: persistent tier-0 waves per compute unit (default: the
    library's, 16); a launch choice like the others, never the result.
    """
    n = int(points.shape[0])
    if clouds is not None:
        clouds = np.ascontiguousarray(clouds, dtype=np.int64)
        n_clouds = int(clouds.shape[0]) - 1
        c_off = ctypes.c_void_p(clouds.ctypes.data)
        if initial_guesses is not None:
            raise ValueError("initial guesses are not supported for batches of clouds")
    dev = points.device
    i32 = dict(dtype=torch.int32, device=dev)
    if n == 0:  # the reference launches a zero-size grid here; return the empty diagram
        return torch.empty(0, **i32), torch.zeros(1, **i32), torch.empty(0, **i32)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev)
        sp = ctypes.c_void_p(stream.cuda_stream)
        if debug:
            torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        if clouds is None:
            ws_bytes = lib.u64("pg_index_workspace_bytes", n)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
            lib.call("pg_build_index", n, _p(points), _p(weights), int(leaf_size), _p(ws), ws_bytes, sp)
        else:
            ws_bytes = lib.u64("pg_batch_workspace_bytes", n, n_clouds)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
            lib.call("pg_build_batch_index", n, n_clouds, _p(points), _p(weights), c_off, int(leaf_size), _p(ws),
                     ws_bytes, sp)
        if radius_model is not None:
            from . import learned

            init_h = learned.predict_on_index(lib, radius_model, ws, n, dev, sp, 0 if clouds is None else n_clouds)
        if debug:
            torch.cuda.synchronize(dev)
            print("BVH Creation time: %f ms" % ((time.perf_counter() - t0) * 1e3))
            if weights is not None:  # as laguerre_ultra.cu:259-265
                print_leaf_statistics(lib, n, ws, sp, dev)
        t1 = time.perf_counter()

        waves = ctypes.c_int32(0)
        lib.call("pg_default_waves", ctypes.byref(waves))
        waves = int(waves.value)
        if waves_per_cu is not None:
            waves = max(4, waves // 16 * int(waves_per_cu))
        stats = torch.zeros(NUM_STATS, dtype=torch.int64, device=dev)
        counts = torch.empty(n, **i32)
        status = torch.empty(n, **i32)
        row_offsets = torch.empty(n, dtype=torch.int64, device=dev)

        guesses = initial_guesses
        goffs = initial_guesses_offsets
        guess_stride = 0
        guess_d2 = None
        seed_window = 0
        if guesses is None and knn > 0:
            # The reference pre-clips each cell by its knn nearest neighbours
            # (voronoi_ultra.cu:70-85, guesses replace it, :158-169).  Here
            # the 16 nearest (the largest count the kNN pass keeps) measured
            # fastest on the 1M uniform and 2M power clouds.
            guess_stride = KNN_PRECLIP if preclip is None else int(preclip)
            guesses = torch.empty(n * guess_stride, **i32)
            guess_d2 = torch.empty(n, dtype=torch.float32, device=dev)
            lib.call("pg_knn_guesses", n, _p(ws), 0 if clouds is None else n_clouds, guess_stride, _p(guesses),
                     _p(guess_d2), sp)
            goffs = None
        if initial_guesses is not None and initial_guesses_offsets is not None:
            n_guess = guesses.numel()
            if n_guess == 0:
                # an empty guess list (all offsets 0) offers nothing, as in the
                # reference (voronoi_ultra.cu:158-169, knn forced to 0); a
                # one-entry placeholder keeps the pointer non-null
                guesses = torch.full((1,), -1, **i32)
            # the kernel indexes guesses through the offsets: reject malformed CSR
            ok = bool(((goffs[1:] >= goffs[:-1]).all() & (goffs[0] >= 0) & (goffs[-1] <= n_guess)).item())
            if not ok:
                raise ValueError("initial_guesses_offsets must be non-decreasing and index into initial_guesses")

        total_blocks = (n + CELL_BLOCK - 1) // CELL_BLOCK
        full_bytes = (4 * lib.u64("pg_row_pool_elems", total_blocks * CELL_BLOCK)
                      + lib.u64("pg_cell_scratch_bytes", n))
        free_bytes, _ = torch.cuda.mem_get_info(dev)
        plan = plan_chunks(total_blocks, CELL_BLOCK, full_bytes, free_bytes)
        chunk_cells = min(plan.chunk_threads, n)
        scratch_bytes = lib.u64("pg_cell_scratch_bytes", chunk_cells)
        scratch = torch.empty(scratch_bytes, dtype=torch.uint8, device=dev)
        pool = [torch.empty(lib.u64("pg_row_pool_elems", chunk_cells) if row_pool is None else max(int(row_pool), 1),
                            **i32)]
        if debug:  # voronoi_ultra.cu:274-281 (here the row pool and the cell scratch)
            print("ConvexCell Memory: %.2f MB/chunk for %d threads, %d chunk(s) over %d points"
                  % ((pool[0].numel() * 4 + scratch_bytes) / 2**20, chunk_cells, plan.num_chunks, n))

        def cells(begin, end):
            rows = pool[0]
            ev = None
            if instrument is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(stream)
            cyc = _p(instrument.cycles if instrument is not None else None)
            if clouds is None:
                lib.call("pg_compute_cells_ex", n, _p(ws), begin, end, _p(guesses), _p(goffs), 0, guess_stride,
                         _p(guess_d2), seed_window, _p(init_h), _p(extent), _p(rows), rows.numel(), _p(row_offsets),
                         _p(counts), _p(status), _p(stats), _p(scratch), scratch_bytes, waves, cyc, sp)
            else:
                lib.call("pg_compute_cells_batch", n, _p(ws), n_clouds, c_off, begin, end, _p(guesses), _p(goffs),
                         guess_stride, _p(guess_d2), _p(rows), rows.numel(), _p(row_offsets), _p(counts), _p(status),
                         _p(stats), _p(scratch), scratch_bytes, waves, cyc, sp)
            if ev is not None:
                ev[1].record(stream)
                instrument.cell_events.append(ev)

        def grow(need):
            """The rows of a launch did not fit the row pool: a pool of the size
            the launch reported, for running it again."""
            pool[0] = None
            pool[0] = torch.empty(need + need // 8 + 65536, **i32)
            if debug:
                print("Row pool grown to %.2f MB (a launch needed %d entries)" % (pool[0].numel() * 4 / 2**20, need))

        scan_tmp = torch.empty(max(lib.u64("pg_scan_tmp_bytes", n + 1), 4), dtype=torch.uint8, device=dev)
        offsets = torch.empty(n + 1, **i32)
        stride = 0
        if plan.num_chunks == 1:
            while True:
                snap = stats.clone()
                cells(0, n)
                src, src_off = pool[0], row_offsets
                lib.call("pg_scan_counts", n, _p(counts), _p(offsets), _p(scan_tmp), sp)
                lib.call("pg_check_symmetry", n, _p(ws), _p(counts), _p(src), _p(src_off), stride, _p(offsets),
                         _p(stats), sp)
                st = stats.cpu()  # the single device-to-host read: E, the symmetry and pool counters
                if int(st[STAT_POOL_SHORT]) == 0:
                    break
                stats.copy_(snap)
                grow(int(st[STAT_POOL_DEMAND]))
        else:
            src_off = torch.empty(n, dtype=torch.int64, device=dev)
            parts, base = [], 0
            for c in range(plan.num_chunks):
                begin = c * chunk_cells
                if begin >= n:
                    break
                end = min(n, begin + chunk_cells)
                m = end - begin
                while True:
                    snap = stats.clone()
                    cells(begin, end)
                    cc = torch.empty(m, **i32)
                    lib.call("pg_chunk_counts", n, _p(ws), begin, end, _p(counts), _p(cc), sp)
                    loc = torch.empty(m + 1, **i32)
                    lib.call("pg_scan_counts", m, _p(cc), _p(loc), _p(scan_tmp), sp)
                    # chunked mode: one read per chunk (its row total and pool counters)
                    vals = torch.cat([loc[m:m + 1].to(torch.int64), stats[STAT_POOL_SHORT:STAT_POOL_DEMAND + 1]]).cpu()
                    if int(vals[1]) == 0:
                        break
                    stats.copy_(snap)
                    grow(int(vals[2]))
                total = int(vals[0])
                packed = torch.empty(max(total, 1), **i32)
                lib.call("pg_pack_rows", n, _p(ws), begin, end, _p(counts), _p(loc), _p(pool[0]), _p(row_offsets),
                         _p(packed), sp)
                src_off[begin:end] = loc[:m].to(torch.int64) + base
                parts.append(packed[:total])
                base += total
            src = torch.cat(parts) if base > 0 else torch.empty(1, **i32)
            lib.call("pg_scan_counts", n, _p(counts), _p(offsets), _p(scan_tmp), sp)
            lib.call("pg_check_symmetry", n, _p(ws), _p(counts), _p(src), _p(src_off), stride, _p(offsets),
                     _p(stats), sp)
            st = stats.cpu()  # E and the symmetry counters
        e_total = int(st[15])
        adjacency = torch.empty(max(e_total, 1), **i32)
        lib.call("pg_emit_csr", n, _p(ws), _p(counts), _p(offsets), _p(src), _p(src_off), stride,
                 _p(adjacency), sp)
        adjacency = adjacency[:e_total]
        if int(st[8]) or int(st[9]):
            adjacency, offsets = _union(lib, n, counts, offsets, adjacency, scan_tmp, sp, dev)
        if debug:
            torch.cuda.synchronize(dev)
            name = "Laguerre" if weights is not None else "Voronoi"
            print("%s Construction time: %f ms" % (name, (time.perf_counter() - t1) * 1e3))
        if instrument is not None:
            instrument.stats = {STAT_NAMES[k]: int(v) for k, v in enumerate(st.tolist()) if k in STAT_NAMES}
            instrument.stats["exact_cells"] = [int(v) - 1 for v in st[STAT_EXACT_LOG:STAT_EXACT_LOG + 7].tolist()
                                               if v > 0]
            instrument.stats["rebuilt_cells"] = [int(v) - 1 for v in
                                                 st[STAT_TENTATIVE_LOG:STAT_TENTATIVE_LOG + 4].tolist() if v > 0]
            instrument.plan = plan
        return adjacency, offsets, status


def leaf_statistics(lib: HipLibrary, n: int, ws: torch.Tensor, sp, dev):
    """(leaves, min, max, total points) of the index's leaves."""
    out = torch.empty(4, dtype=torch.int64, device=dev)
    lib.call("pg_leaf_stats", n, _p(ws), _p(out), sp)
    return tuple(int(v) for v in out.cpu().tolist())


def print_leaf_statistics(lib: HipLibrary, n: int, ws: torch.Tensor, sp, dev):
    """The debug print of pwr_bvh.cu:444-449."""
    leaves, lo, hi, total = leaf_statistics(lib, n, ws, sp, dev)
    if leaves == 0:
        print("BVH Leaf Statistics: No leaves found")
        return
    print("BVH Leaf Statistics:")
    print("  Total leaves: %d" % leaves)
    print("  Min points per leaf: %d" % lo)
    print("  Max points per leaf: %d" % hi)
    print("  Mean points per leaf: %.2f" % (total / leaves))
    print("  Total points in leaves: %d" % total)


def knn(lib: HipLibrary, points: torch.Tensor, k: int = 8, leaf_size: int = -1):
    """``(ids int32 [N, k], d2 float32 [N, k])``: the k nearest neighbours of each
    point (itself excluded), as the reference's kNN pass (pwr_bvh.cu:159-215)."""
    n = int(points.shape[0])
    dev = points.device
    ids = torch.empty((n, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((n, k), dtype=torch.float32, device=dev)
    if n == 0:
        return ids, d2
    with torch.cuda.device(dev):
        sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        ws_bytes = lib.u64("pg_index_workspace_bytes", n)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        lib.call("pg_build_index", n, _p(points), None, int(leaf_size), _p(ws), ws_bytes, sp)
        lib.call("pg_knn", n, _p(ws), int(k), _p(ids), _p(d2), sp)
    return ids, d2


def split_batch(adjacency: torch.Tensor, offsets: torch.Tensor, status: torch.Tensor, clouds, cloud_entries=None,
                lib: HipLibrary | None = None):
    """Per-cloud ``(adjacency, offsets, status)`` of a batch diagram, with
    cloud-local ids; views into two new device arrays written by one
    ``pg_batch_localize`` pass.  ``cloud_entries`` (host, optional): each
    cloud's adjacency entry count, else read once."""
    if lib is None:
        from ._backend import _C

        if _C is None:
            raise RuntimeError("libparagram_hip.so is not loaded")
        lib = _C.lib
    dev = offsets.device
    clouds = np.ascontiguousarray(clouds, dtype=np.int64)
    b = clouds.shape[0] - 1
    n = int(offsets.numel()) - 1
    e_total = int(adjacency.numel())
    i32 = dict(dtype=torch.int32, device=dev)
    cl_dev = torch.from_numpy(clouds).to(dev)
    local_adj = torch.empty(max(e_total, 1), **i32)
    local_off = torch.empty(n + b, **i32)
    adj_in = adjacency if e_total > 0 else torch.zeros(1, **i32)
    with torch.cuda.device(dev):
        sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        lib.call("pg_batch_localize", n, b, _p(cl_dev), _p(offsets), _p(adj_in), _p(local_adj), _p(local_off), sp)
    # cloud c's offsets: local_off[firsts[c] .. firsts[c] + n_c] (n_c + 1 entries)
    firsts = clouds[:-1] + np.arange(b, dtype=np.int64)
    if cloud_entries is None:  # adjacency slices need the per-cloud entry counts on the host (one read)
        bounds = offsets[cl_dev].to(torch.int64).cpu().numpy()
        seg_host = bounds[1:] - bounds[:-1]
    else:
        seg_host = np.asarray(cloud_entries)
    res = []
    e0 = 0
    for c in range(b):
        a, z = int(clouds[c]), int(clouds[c + 1])
        m = int(seg_host[c])
        res.append((local_adj[e0:e0 + m], local_off[int(firsts[c]):int(firsts[c]) + (z - a) + 1], status[a:z]))
        e0 += m
    return res


def _union(lib, n, counts, offsets, adjacency, scan_tmp, sp, dev):
    """Union symmetrisation (rare path: only when some edge lacks its reverse)."""
    i32 = dict(dtype=torch.int32, device=dev)
    keep = torch.empty(n, **i32)
    extra = torch.empty(n, **i32)
    new_counts = torch.empty(n, **i32)
    new_offsets = torch.empty(n + 1, **i32)
    lib.call("pg_fill_u32", _p(extra), n, 0, sp)
    lib.call("pg_symmetrize_union", n, _p(counts), _p(offsets), _p(adjacency), _p(keep), _p(extra), _p(new_counts),
             None, None, 0, sp)
    lib.call("pg_symmetrize_union", n, _p(counts), _p(offsets), _p(adjacency), _p(keep), _p(extra), _p(new_counts),
             None, None, 1, sp)
    lib.call("pg_scan_counts", n, _p(new_counts), _p(new_offsets), _p(scan_tmp), sp)
    e2 = int(new_offsets[n].item())
    new_adj = torch.empty(max(e2, 1), **i32)
    lib.call("pg_fill_u32", _p(extra), n, 0, sp)
    lib.call("pg_symmetrize_union", n, _p(counts), _p(offsets), _p(adjacency), _p(keep), _p(extra), _p(new_counts),
             _p(new_offsets), _p(new_adj), 2, sp)
    lib.call("pg_symmetrize_union", n, _p(counts), _p(offsets), _p(adjacency), _p(keep), _p(extra), _p(new_counts),
             _p(new_offsets), _p(new_adj), 3, sp)
    return new_adj[:e2], new_offsets
