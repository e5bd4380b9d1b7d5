"""The host-core build (paragram_amd.cpu, csrc/pg_cpu.cpp: the clipping
algorithm of the GPU kernels in C++/OpenMP) against the oracle, bit-exact:
uniform and power clouds with hidden cells, duplicates, tiny inputs, the
golden fixtures, the cubic lattice (face adjacency, as the GPU path), and
the reference's two callables on CPU tensors."""
import os

import numpy as np
import pytest

import oracle
from paragram_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def host():
    from paragram_amd import cpu

    return cpu.CpuLibrary()


def same_as_oracle(host, pts, w=None):
    adj, off, st, stats = host.diagram(pts, w)
    a2, o2, s2, _ = oracle.diagram(pts, w)
    assert np.array_equal(st, s2), f"status differs in {int((st != s2).sum())} cells"
    assert np.array_equal(off, o2) and np.array_equal(adj, a2)
    return stats


def test_c1_voronoi_10k(host):
    stats = same_as_oracle(host, workloads.uniform(10_000, 0))
    assert stats["unresolved"] == 0 and stats["clips"] > 0


def test_power_hidden_and_clustered(host):
    rng = np.random.default_rng(5)
    pts = rng.random((5000, 3)).astype(np.float32)
    w = rng.uniform(0, 4e-3, 5000).astype(np.float32)
    same_as_oracle(host, pts, w)
    same_as_oracle(host, *workloads.clustered_power(20_000, 2))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 17, 33, 200])
def test_small_sizes(host, n):
    same_as_oracle(host, workloads.uniform(n, 100 + n))


def test_duplicates_and_lidar(host):
    pts = workloads.uniform(1000, 9)
    same_as_oracle(host, np.concatenate([pts, pts[:50]], 0))
    same_as_oracle(host, workloads.lidar_batch(3, max_clouds=1)[0][:30_000])


@pytest.mark.parametrize("name", ["voronoi_uniform_300", "power_hidden_400", "power_clustered_500", "lattice_4"])
def test_golden(host, name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        g = {k: f[k] for k in f.files}
    adj, off, st, _ = host.diagram(g["points"], g.get("weights"))
    assert np.array_equal(adj, g["adjacency"]) and np.array_equal(off, g["offsets"])
    assert np.array_equal(st, g["status"])


def test_backend_callables_on_cpu_tensors(host):
    import torch

    from paragram_amd import cpu

    be = cpu.CpuBackend(host)
    pts = workloads.uniform(3000, 4)
    a, o, s = be.build_voronoi_ultra(torch.from_numpy(pts))
    assert a.dtype == torch.int32 and o.shape == (3001,) and s.shape == (3000,)
    a2, o2, s2, _ = oracle.diagram(pts)
    assert np.array_equal(a.numpy(), a2) and np.array_equal(o.numpy(), o2)
    w = np.random.default_rng(1).uniform(0, 1e-3, 3000).astype(np.float32)
    a, o, s = be.build_laguerre_voronoi_ultra(torch.from_numpy(pts), torch.from_numpy(w).reshape(-1, 1))
    a2, o2, s2, _ = oracle.diagram(pts, w)
    assert np.array_equal(a.numpy(), a2) and np.array_equal(s.numpy(), s2)
