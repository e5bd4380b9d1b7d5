"""The LBVH (pg_build_index) against brute force: the Morton-sorted points
are a permutation of the input, leaves partition the sorted points (every
primitive exactly once, leaves of at most leaf_size points), every child box
is the exact FP32 AABB of its subtree's points (so it contains them and sits
inside its parent's), and every max weight is the subtree's maximum
(replaces the cuBQL builder and the refit of pwr_bvh.cu:42-157,217-339).
Also the debug output of a power diagram (leaf statistics, as
laguerre_ultra.cu:259-265 / pwr_bvh.cu:444-449)."""
import ctypes

import numpy as np
import pytest

from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def build(gpu, pts, w, leaf_size):
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    lib = _C.lib
    n = pts.shape[0]
    P = torch.from_numpy(pts).to(gpu)
    W = None if w is None else torch.from_numpy(w).to(gpu)
    ws_bytes = lib.u64("pg_index_workspace_bytes", n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=gpu)
    sp = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    lib.call("pg_build_index", n, _engine._p(P), _engine._p(W), leaf_size, _engine._p(ws), ws_bytes, sp)
    offs = (ctypes.c_uint64 * 6)()
    lib.call("pg_index_layout", n, offs)
    stats = _engine.leaf_statistics(lib, n, ws, sp, gpu)
    return ws.cpu().numpy(), list(offs), stats


def check_hierarchy(pts, w, leaf_size, blob, offs):
    n = pts.shape[0]
    hdr = blob[offs[0]:offs[0] + 72]
    root = int(hdr[48:52].view(np.uint32)[0])
    sorted4 = blob[offs[1]:offs[1] + 16 * n].view(np.float32).reshape(n, 4)
    perm = blob[offs[2]:offs[2] + 4 * n].view(np.uint32).astype(np.int64)
    inv = blob[offs[3]:offs[3] + 4 * n].view(np.uint32).astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n)), "perm is not a permutation"
    assert np.array_equal(inv[perm], np.arange(n))
    ww = np.zeros(n, np.float32) if w is None else w
    assert np.array_equal(sorted4[:, :3], pts[perm]) and np.array_equal(sorted4[:, 3], ww[perm])
    ni = max(n - 1, 1)
    nodes = blob[offs[4]:offs[4] + 64 * ni].view(np.float32).reshape(ni, 16)
    refs = nodes.view(np.uint32)

    def is_leaf(r):
        return r & 0x80000000

    def leaf_range(r):
        first = (r & 0x7FFFFFFF) >> 5
        return first, first + (r & 31) + 1

    covered = np.zeros(n, np.int64)
    leaves = []
    # iterative post-order: the range of every internal node
    rng = {}
    stack = [(root, False)] if not is_leaf(root) else []
    if is_leaf(root):
        a, b = leaf_range(root)
        covered[a:b] += 1
        leaves.append(b - a)
    while stack:
        node, done = stack.pop()
        kids = (int(refs[node, 7]), int(refs[node, 15]))
        if not done:
            stack.append((node, True))
            for r in kids:
                if not is_leaf(r):
                    stack.append((r, False))
            continue
        lo_hi = []
        for c, r in enumerate(kids):
            if is_leaf(r):
                a, b = leaf_range(r)
                covered[a:b] += 1
                leaves.append(b - a)
            else:
                a, b = rng[r]
            lo_hi.append((a, b))
            box_lo, box_hi = nodes[node, 8 * c:8 * c + 3], nodes[node, 8 * c + 4:8 * c + 7]
            sub = sorted4[a:b]
            assert np.array_equal(box_lo, sub[:, :3].min(0)), f"node {node} child {c}: box lo is not the subtree AABB"
            assert np.array_equal(box_hi, sub[:, :3].max(0)), f"node {node} child {c}: box hi is not the subtree AABB"
            assert nodes[node, 8 * c + 3] == sub[:, 3].max(), f"node {node} child {c}: max weight"
        (a0, b0), (a1, b1) = lo_hi
        assert b0 == a1, f"node {node}: children ranges are not adjacent"
        rng[node] = (a0, b1)
    assert np.all(covered == 1), "a primitive is in no leaf or in two"
    assert max(leaves) <= leaf_size
    return len(leaves), min(leaves), max(leaves), sum(leaves)


@pytest.mark.parametrize("n,leaf", [(1, 32), (2, 32), (33, 4), (5000, 1), (100_000, 32), (100_000, 10)])
def test_lbvh_invariants_against_brute_force(gpu, n, leaf):
    log(f"LBVH invariants: n={n}, leaf_size={leaf}")
    pts, w = workloads.clustered_power(n, 7) if n >= 1000 else (workloads.uniform(n, 7), None)
    blob, offs, stats = build(gpu, pts, w, leaf)
    want = check_hierarchy(pts, w, leaf, blob, offs)
    assert stats == want, (stats, want)


def test_debug_output_of_power_diagram(gpu, capsys):
    import torch

    import paragram_amd

    pts, w = workloads.clustered_power(20_000, 3)
    d = paragram_amd.power_diagram(torch.from_numpy(pts).to(gpu), torch.from_numpy(w).to(gpu), debug=True)
    torch.cuda.synchronize()
    out = capsys.readouterr().out
    for line in ("BVH Creation time:", "BVH Leaf Statistics:", "  Total leaves:", "  Min points per leaf:",
                 "  Max points per leaf:", "  Mean points per leaf:", "  Total points in leaves: 20000",
                 "ConvexCell Memory:", "Laguerre Construction time:"):
        assert line in out, line
    assert d.offsets.shape[0] == 20_001
