"""pg_compute_cells_noalloc (cell construction with the library's static cell
scratch, for callers that cannot allocate) against pg_compute_cells_ex with a
caller-allocated scratch, through the C ABI: two calls in a row on the same
GPU, the second with a smaller input (the static array keeps the first
call's contents), each checked byte for byte (CSR and statuses) and in its
work statistics."""
import ctypes

import numpy as np
import pytest

from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def diagram(lib, dev, pts, w, noalloc):
    """Index, kNN pre-clip, cells (through either entry point), offsets and
    CSR emission, all through the C ABI; returns numpy (adjacency, offsets,
    status, stats)."""
    import torch

    from paragram_amd import _engine

    p = _engine._p
    n = pts.shape[0]
    i32 = dict(dtype=torch.int32, device=dev)
    P = torch.from_numpy(pts).to(dev)
    W = None if w is None else torch.from_numpy(w).to(dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    ws_bytes = lib.u64("pg_index_workspace_bytes", n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    lib.call("pg_build_index", n, p(P), p(W), -1, p(ws), ws_bytes, sp)
    k = _engine.KNN_PRECLIP
    guesses = torch.empty(n * k, **i32)
    d2 = torch.empty(n, dtype=torch.float32, device=dev)
    lib.call("pg_knn_guesses", n, p(ws), 0, k, p(guesses), p(d2), sp)
    waves = ctypes.c_int32(0)
    lib.call("pg_default_waves", ctypes.byref(waves))
    rows = torch.empty(lib.u64("pg_row_pool_elems", n), **i32)
    row_offsets = torch.empty(n, dtype=torch.int64, device=dev)
    counts = torch.empty(n, **i32)
    status = torch.empty(n, **i32)
    stats = torch.zeros(_engine.NUM_STATS, dtype=torch.int64, device=dev)
    if noalloc:
        lib.call("pg_compute_cells_noalloc", n, p(ws), 0, n, p(guesses), None, 0, k, p(d2), 0, None, None, p(rows),
                 rows.numel(), p(row_offsets), p(counts), p(status), p(stats), waves.value, None, sp)
    else:
        sbytes = lib.u64("pg_cell_scratch_bytes", n)
        scratch = torch.empty(sbytes, dtype=torch.uint8, device=dev)
        lib.call("pg_compute_cells_ex", n, p(ws), 0, n, p(guesses), None, 0, k, p(d2), 0, None, None, p(rows),
                 rows.numel(), p(row_offsets), p(counts), p(status), p(stats), p(scratch), sbytes, waves.value,
                 None, sp)
    offsets = torch.empty(n + 1, **i32)
    tmp = torch.empty(max(lib.u64("pg_scan_tmp_bytes", n + 1), 4), dtype=torch.uint8, device=dev)
    lib.call("pg_scan_counts", n, p(counts), p(offsets), p(tmp), sp)
    lib.call("pg_check_symmetry", n, p(ws), p(counts), p(rows), p(row_offsets), 0, p(offsets), p(stats), sp)
    st = stats.cpu().numpy()
    e = int(st[15])
    adj = torch.empty(max(e, 1), **i32)
    lib.call("pg_emit_csr", n, p(ws), p(counts), p(offsets), p(rows), p(row_offsets), 0, p(adj), sp)
    torch.cuda.synchronize(dev)
    return adj[:e].cpu().numpy(), offsets.cpu().numpy(), status.cpu().numpy(), st


def test_noalloc_entry_point_matches_allocating_one(gpu):
    from paragram_amd._backend import _C

    lib = _C.lib
    cap = ctypes.c_int64(0)
    lib.call("pg_noalloc_max_cells", ctypes.byref(cap))
    limit = 48 << 20  # the static cell scratch: the largest range whose scratch fits
    assert lib.u64("pg_cell_scratch_bytes", cap.value) <= limit < lib.u64("pg_cell_scratch_bytes", cap.value + 1)
    assert cap.value == 1_571_837
    rng = np.random.default_rng(5)
    big = workloads.uniform(300_000, 41)
    small_pts = rng.random((20_000, 3)).astype(np.float32)
    small_w = rng.uniform(0, 1e-4, 20_000).astype(np.float32)
    for name, pts, w in (("300k uniform", big, None), ("20k power (smaller, second)", small_pts, small_w)):
        log(f"noalloc vs allocating entry point: {name}")
        want = diagram(lib, gpu, pts, w, noalloc=False)
        got = diagram(lib, gpu, pts, w, noalloc=True)
        for label, x, y in zip(("adjacency", "offsets", "status"), got[:3], want[:3]):
            assert np.array_equal(x, y), f"{name}: {label} differs"
        for slot in (0, 3, 4, 5, 6, 7, 10, 12, 13, 15):  # the work and outcome counters agree too
            assert got[3][slot] == want[3][slot], (name, slot)
        assert got[3][12] == 0 and got[3][13] == 0
    # a range beyond the static scratch is refused before any device work
    rc = lib.lib.pg_compute_cells_noalloc(cap.value + 1, None, 0, cap.value + 1, None, None, 0, 0, None, 0, None,
                                          None, None, 0, None, None, None, None, 1, None, None)
    assert rc == 1 and b"too large for the static cell scratch" in lib.lib.pg_last_error()
