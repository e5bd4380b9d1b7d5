"""GPU checks of the CSR stage through the C ABI: the union symmetrisation
(pg_symmetrize_union, phases 0-3) on hand-built asymmetric rows against the
oracle's union (adjacency.cuh:78-159 semantics: union of u->v and v->u,
entries pointing at empty cells dropped), and empty initial-guess lists."""
import ctypes

import numpy as np
import pytest

import oracle
from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def random_rows(rng, n, max_deg, empty_frac):
    """Sorted rows without self loops; some cells empty (count 0) yet listed by others."""
    counts = rng.integers(1, max_deg + 1, n)
    counts[rng.random(n) < empty_frac] = 0
    rows = []
    for u in range(n):
        cand = rng.choice(n - 1, counts[u], replace=False)
        cand = np.where(cand >= u, cand + 1, cand)  # skip u itself
        rows.append(np.sort(cand).astype(np.int32))
    return counts.astype(np.int32), rows


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_union_symmetrization_matches_oracle(gpu, seed):
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    lib = _C.lib
    rng = np.random.default_rng(seed)
    n = [500, 3000, 20000][seed]
    counts, rows = random_rows(rng, n, 12, 0.05)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    flat = np.concatenate(rows).astype(np.int32)
    log(f"union symmetrisation: n={n}, {flat.size} directed entries, {(counts == 0).sum()} empty cells")
    ref = oracle.Rows(counts, np.zeros(n, np.int32), flat, {})
    want_adj, want_off = oracle.symmetrize(ref)
    assert want_adj.size > flat.size * 0.9  # the union adds the missing reverse edges

    dev = torch.device(gpu)
    c, o, a = (torch.from_numpy(x).to(dev) for x in (counts, offsets, flat))
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    scan_tmp = torch.empty(max(lib.u64("pg_scan_tmp_bytes", n + 1), 4), dtype=torch.uint8, device=dev)
    adj, off = _engine._union(lib, n, c, o, a, scan_tmp, sp, dev)
    assert np.array_equal(off.cpu().numpy(), want_off)
    assert np.array_equal(adj.cpu().numpy(), want_adj)
    log("union symmetrisation: outputs equal the oracle's")


def test_empty_initial_guesses_accepted(gpu):
    """A zero-length guess list with all-zero offsets offers nothing (the
    reference then skips the pre-clip, knn forced to 0): the diagram equals
    the plain call's."""
    import torch

    import paragram_amd

    log("empty initial_guesses with zero offsets")
    pts = workloads.uniform(2000, 13)
    p = torch.from_numpy(pts).to(gpu)
    g = torch.empty(0, dtype=torch.int32, device=gpu)
    offs = torch.zeros(2001, dtype=torch.int32, device=gpu)
    d = paragram_amd.voronoi_diagram(p, initial_guesses=g, initial_guesses_offsets=offs)
    e = paragram_amd.voronoi_diagram(p)
    for x, y in ((d.adjacency, e.adjacency), (d.offsets, e.offsets), (d.status, e.status)):
        assert torch.equal(x, y)
    adj, off, st, _ = oracle.diagram(pts)
    assert np.array_equal(d.adjacency.cpu().numpy(), adj) and np.array_equal(d.status.cpu().numpy(), st)
