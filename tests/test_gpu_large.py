"""Beyond the reference's int32 index limit: 2^25 + 1000 points in one call.

The reference computes the neighbour slot of a cell as ``idx * 64`` in int32
(voronoi_ultra.cu:135, laguerre_ultra.cu:167), which overflows for
N >= 2^25 = 33,554,432; this build's limit is 2^26 - 1 points (leaf references
keep 26 bits of a first index).  The diagram of a uniform cloud just past 2^25
is checked through size-independent CSR invariants on the GPU (offsets well
formed, ids in range, rows strictly increasing without self loops, every edge
with its reverse) and against the oracle on cells on both sides of 2^25 and in
the last leaf of the index.
"""
import numpy as np
import pytest
import torch

import oracle
from paragram_amd import workloads

pytestmark = pytest.mark.gpu

N_LARGE = (1 << 25) + 1000


def gpu_csr_invariants(adj: torch.Tensor, off: torch.Tensor, status: torch.Tensor) -> None:
    n = status.numel()
    off64 = off.to(torch.int64)
    e = int(off64[-1].item())
    assert int(off64[0].item()) == 0 and e == adj.numel()
    deg = off64[1:] - off64[:-1]
    assert bool((deg >= 0).all())
    a = adj.to(torch.int64)
    assert bool(((a >= 0) & (a < n)).all())
    rows = torch.repeat_interleave(torch.arange(n, device=adj.device), deg, output_size=e)
    assert not bool((a == rows).any()), "self loop"
    keys = rows * n + a
    del rows
    assert bool((keys[1:] > keys[:-1]).all()), "a row is not strictly increasing"
    rev = torch.sort(a * n + torch.repeat_interleave(torch.arange(n, device=adj.device), deg, output_size=e)).values
    assert torch.equal(rev, keys), "an edge has no reverse"


def test_beyond_int32_slot_index(gpu):
    import paragram_amd

    pts_np = workloads.uniform(N_LARGE, 7)
    print(f"uniform Voronoi, {N_LARGE} points (seed 7)", flush=True)
    pts = torch.from_numpy(pts_np).to(gpu)
    d = paragram_amd.voronoi_diagram(pts)
    torch.cuda.synchronize()
    assert d.offsets.numel() == N_LARGE + 1 and d.status.numel() == N_LARGE
    assert bool((d.status == 0).all())
    gpu_csr_invariants(d.adjacency, d.offsets, d.status)
    mean_faces = d.adjacency.numel() / N_LARGE
    assert 15.0 < mean_faces < 16.0  # uniform cloud: 15.5 faces per cell
    rng = np.random.default_rng(7)
    edge = 1 << 25
    sample = np.unique(np.concatenate([
        np.arange(edge - 300, edge + 300), np.arange(N_LARGE - 200, N_LARGE), rng.choice(N_LARGE, 500, replace=False)]))
    off = d.offsets.cpu().numpy()
    status = d.status.cpu().numpy()
    ref = oracle.cells(pts_np, None, sample)
    for k, i in enumerate(sample):
        lo, hi = int(off[i]), int(off[i + 1])
        row = d.adjacency[lo:hi].cpu().numpy()
        assert status[i] == ref.status[k] and np.array_equal(row, ref.row(k)), i
