"""GPU parity of the batched launch: every cloud of a batch must equal the
single-cloud call on that cloud (and, for a small batch, the oracle), byte for
byte: adjacency with cloud-local ids, offsets and status."""
import numpy as np
import pytest

import oracle
from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def single(gpu, pts, w=None):
    import torch

    import paragram_amd

    p = torch.from_numpy(np.ascontiguousarray(pts)).to(gpu)
    d = (paragram_amd.voronoi_diagram(p) if w is None
         else paragram_amd.power_diagram(p, torch.from_numpy(np.ascontiguousarray(w)).to(gpu)))
    return d.adjacency.cpu().numpy(), d.offsets.cpu().numpy(), d.status.cpu().numpy()


def batch(gpu, clouds, weights=None):
    import torch

    import paragram_amd

    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in clouds])]).astype(np.int64)
    p = torch.from_numpy(np.ascontiguousarray(np.concatenate(clouds).reshape(-1, 3))).to(gpu)
    if weights is None:
        ds = paragram_amd.batched_voronoi_diagram(p, off)
    else:
        w = torch.from_numpy(np.concatenate(weights).astype(np.float32)).to(gpu)
        ds = paragram_amd.batched_power_diagram(p, w, off)
    assert len(ds) == len(clouds)
    return [(d.adjacency.cpu().numpy(), d.offsets.cpu().numpy(), d.status.cpu().numpy()) for d in ds]


def assert_same(got, want, what):
    for name, g, w in zip(("adjacency", "offsets", "status"), got, want):
        assert np.array_equal(g, w), f"{what}: {name} differs"


def test_batch_lidar_equals_single_calls(gpu):
    log("batch of 8 LiDAR clouds vs single calls")
    clouds = workloads.lidar_batch(3, max_clouds=8)
    got = batch(gpu, clouds)
    for k, c in enumerate(clouds):
        assert_same(got[k], single(gpu, c), f"cloud {k} ({c.shape[0]} points)")


def test_batch_ragged_sizes(gpu):
    sizes = [0, 1, 2, 5, 16, 17, 100, 0, 3000, 1, 64]
    log(f"batch of ragged uniform clouds {sizes}")
    rng = np.random.default_rng(11)
    clouds = [(rng.random((m, 3)) * rng.uniform(0.1, 10.0) + rng.uniform(-5, 5, 3)).astype(np.float32)
              for m in sizes]
    got = batch(gpu, clouds)
    for k, c in enumerate(clouds):
        if c.shape[0] == 0:
            assert got[k][0].size == 0 and list(got[k][1]) == [0] and got[k][2].size == 0
            continue
        assert_same(got[k], single(gpu, c), f"cloud {k} ({c.shape[0]} points)")


def test_batch_matches_oracle(gpu):
    log("batch of 5 small clouds vs the oracle")
    rng = np.random.default_rng(12)
    clouds = [rng.random((m, 3)).astype(np.float32) for m in (300, 1, 2000, 40, 777)]
    got = batch(gpu, clouds)
    for k, c in enumerate(clouds):
        adj, off, st, _ = oracle.diagram(c)
        assert_same(got[k], (adj, off, st), f"cloud {k}")


def test_batch_power_equals_single_calls(gpu):
    log("batch of 4 clustered power clouds vs single calls")
    clouds, weights = [], []
    for s in range(4):
        p, w = workloads.clustered_power(5000 + 3000 * s, 20 + s)
        clouds.append(p)
        weights.append(w)
    got = batch(gpu, clouds, weights)
    for k in range(4):
        assert_same(got[k], single(gpu, clouds[k], weights[k]), f"power cloud {k}")


def test_batch_chunked_identical(gpu, monkeypatch):
    log("batch of 6 LiDAR clouds: unchunked vs PARAGRAM_NUM_CHUNKS=3")
    clouds = workloads.lidar_batch(3, max_clouds=6)
    monkeypatch.delenv("PARAGRAM_NUM_CHUNKS", raising=False)
    a = batch(gpu, clouds)
    monkeypatch.setenv("PARAGRAM_NUM_CHUNKS", "3")
    b = batch(gpu, clouds)
    for k in range(len(clouds)):
        assert_same(a[k], b[k], f"cloud {k}")
