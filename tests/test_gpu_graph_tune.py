"""GPU: the graph-captured pipeline and the autotuned entry points return
exactly what the plain calls return."""
import numpy as np
import pytest

from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def eager(gpu, pts, w=None):
    import torch

    import paragram_amd

    P = torch.from_numpy(pts).to(gpu)
    d = paragram_amd.voronoi_diagram(P) if w is None else paragram_amd.power_diagram(P, torch.from_numpy(w).to(gpu))
    return d.adjacency.clone(), d.offsets.clone(), d.status.clone()


@pytest.mark.parametrize("weighted", [False, True])
def test_graph_replays_equal_eager_calls(gpu, weighted):
    import torch

    from paragram_amd.graph import GraphedDiagram

    log(f"graph capture and replays, weighted={weighted}")
    n = 20_000
    g = GraphedDiagram(n, weighted=weighted, device=gpu)
    for seed in (1, 2, 3, 1):
        if weighted:
            pts, w = workloads.clustered_power(n, seed)
        else:
            pts, w = workloads.uniform(n, seed), None
        want = eager(gpu, pts, w)  # eager work between replays
        P = torch.from_numpy(pts).to(gpu)
        W = None if w is None else torch.from_numpy(w).to(gpu)
        got = g(P, W)
        torch.cuda.synchronize()
        for x, y in zip(got, want):
            assert torch.equal(x, y), f"replay for seed {seed} differs"


@pytest.mark.parametrize("weighted", [False, True])
def test_graphed_batch_replays_equal_eager_batches(gpu, weighted):
    """A batch of clouds captured in one graph (index, kNN pre-clip, cells,
    CSR): replays for new coordinates of the same cloud sizes, with eager
    batched calls between them, equal the eager batched calls."""
    import torch

    import paragram_amd
    from paragram_amd.graph import GraphedBatch

    log(f"graphed batch, weighted={weighted}")
    sizes = [3000, 1, 0, 17, 5000, 64, 2]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    g = GraphedBatch(off, weighted=weighted, device=gpu)
    for seed in (4, 5, 4):
        rng = np.random.default_rng(seed)
        pts = rng.random((n, 3)).astype(np.float32)
        w = rng.uniform(0, 1e-4, n).astype(np.float32) if weighted else None
        P = torch.from_numpy(pts).to(gpu)
        W = None if w is None else torch.from_numpy(w).to(gpu)
        want = (paragram_amd.batched_voronoi_diagram(P, off) if w is None
                else paragram_amd.batched_power_diagram(P, W, off))
        got = g(P, W)
        torch.cuda.synchronize()
        assert len(got) == len(want) == len(sizes)
        for c, ((a, o, s), d) in enumerate(zip(got, want)):
            assert torch.equal(a, d.adjacency) and torch.equal(o, d.offsets) and torch.equal(s, d.status), \
                f"replay for seed {seed}: cloud {c} differs"


def test_autotuned_calls_equal_plain(gpu):
    import torch

    import paragram_amd
    from paragram_amd import autotune

    log("autotune on 100k uniform points, then tuned == plain")
    pts = workloads.uniform(100_000, 5)
    P = torch.from_numpy(pts).to(gpu)
    cfg = autotune.tune(P, force=True)
    log(f"tuned: {cfg}")
    assert cfg["leaf_size"] in autotune.LEAF_SIZES and cfg["preclip"] in autotune.PRECLIPS
    assert cfg["waves_per_cu"] in autotune.WAVES_PER_CU
    a = autotune.voronoi_diagram(P)
    b = paragram_amd.voronoi_diagram(P)
    assert torch.equal(a.adjacency, b.adjacency) and torch.equal(a.offsets, b.offsets)
    clouds = workloads.lidar_batch(3, max_clouds=3)
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in clouds])])
    Q = torch.from_numpy(np.concatenate(clouds)).to(gpu)
    ta = autotune.batched_voronoi_diagram(Q, off)
    tb = paragram_amd.batched_voronoi_diagram(Q, off)
    for x, y in zip(ta, tb):
        assert torch.equal(x.adjacency, y.adjacency) and torch.equal(x.status, y.status)
