"""GPU parity: the HIP path against the CPU oracle, cell by cell.

Bar: bit-exact adjacency, offsets and status (integer output).  Every cell is
checked up to 1M points.  Inputs are the seeded generators of
paragram_amd.workloads (configs C1, C2) plus edge cases the reference's code
paths reach (N = 0, 1, 2, below and at the leaf size, duplicates, coplanar
clamped faces, hidden power cells, initial guesses).
"""
import os

import numpy as np
import pytest

import oracle
from paragram_amd import workloads
from tests.helpers import csr_invariants, csr_invariants_fast

pytestmark = pytest.mark.gpu


def run(gpu, points, weights=None, **kw):
    import torch

    import paragram_amd

    pts = torch.from_numpy(np.ascontiguousarray(points)).to(gpu)
    if weights is None:
        d = paragram_amd.voronoi_diagram(pts, **kw)
    else:
        d = paragram_amd.power_diagram(pts, torch.from_numpy(np.ascontiguousarray(weights)).to(gpu), **kw)
    torch.cuda.synchronize()
    return d.adjacency.cpu().numpy(), d.offsets.cpu().numpy(), d.status.cpu().numpy()


def assert_matches_oracle(got, points, weights=None):
    adj, off, st, _ = oracle.diagram(points, weights)
    g_adj, g_off, g_st = got
    assert np.array_equal(g_st, st), f"status differs in {int((g_st != st).sum())} cells"
    assert np.array_equal(g_off, off), f"offsets differ ({int((np.diff(g_off) != np.diff(off)).sum())} rows)"
    assert np.array_equal(g_adj, adj), "adjacency differs"


def log(msg):
    print(msg, flush=True)


def test_c1_voronoi_10k(gpu):
    log("C1 voronoi 10k seed 0")
    pts = workloads.uniform(10_000, 0)
    got = run(gpu, pts)
    csr_invariants(*got)
    assert_matches_oracle(got, pts)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 8, 9, 33, 200])
def test_small_sizes(gpu, n):
    log(f"voronoi small n={n}")
    pts = workloads.uniform(n, 100 + n)
    got = run(gpu, pts)
    if n == 0:
        assert got[0].size == 0 and list(got[1]) == [0] and got[2].size == 0
        return
    csr_invariants(*got)
    assert_matches_oracle(got, pts)


@pytest.mark.parametrize("leaf", [1, 4, 16, 32])
def test_leaf_sizes(gpu, leaf):
    log(f"voronoi leaf_size={leaf}")
    pts = workloads.uniform(3000, 11)
    got = run(gpu, pts, leaf_size=leaf)
    assert_matches_oracle(got, pts)


def test_power_hidden_cells(gpu):
    log("power 5000 with hidden cells")
    rng = np.random.default_rng(5)
    pts = rng.random((5000, 3)).astype(np.float32)
    w = rng.uniform(0, 4e-3, 5000).astype(np.float32)
    got = run(gpu, pts, w)
    csr_invariants(*got)
    assert (got[2] == 3).sum() > 0
    assert_matches_oracle(got, pts, w)


def test_power_zero_weights_equals_voronoi(gpu):
    log("power with zero weights == voronoi")
    pts = workloads.uniform(4000, 3)
    a = run(gpu, pts)
    b = run(gpu, pts, np.zeros(4000, np.float32))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_clustered_clamped_power(gpu):
    log("C3-shaped clustered power 20k")
    pts, w = workloads.clustered_power(20_000, 2)
    got = run(gpu, pts, w)
    csr_invariants(*got)
    assert_matches_oracle(got, pts, w)


def test_duplicates(gpu):
    log("duplicate points")
    pts = workloads.uniform(1000, 9)
    pts = np.concatenate([pts, pts[:50]], 0)
    got = run(gpu, pts)
    csr_invariants(*got)
    assert_matches_oracle(got, pts)


def test_initial_guesses(gpu):
    import torch

    log("initial guesses")
    pts = workloads.uniform(3000, 12)
    n = pts.shape[0]
    rng = np.random.default_rng(0)
    g = rng.integers(-1, n, size=n * 5).astype(np.int32)
    offs = np.arange(0, n * 5 + 1, 5, dtype=np.int32)
    got = run(gpu, pts, initial_guesses=torch.from_numpy(g).to(gpu),
              initial_guesses_offsets=torch.from_numpy(offs).to(gpu))
    assert_matches_oracle(got, pts)


def test_lidar_cloud(gpu):
    log("one LiDAR-like cloud")
    cloud = workloads.lidar_batch(3, max_clouds=1)[0]
    got = run(gpu, cloud)
    csr_invariants(*got)
    assert_matches_oracle(got, cloud)


def test_c2_chunked_identical(gpu, monkeypatch):
    log("C2 1M: unchunked vs PARAGRAM_NUM_CHUNKS=4")
    pts = workloads.uniform(1_000_000, 1)
    monkeypatch.delenv("PARAGRAM_NUM_CHUNKS", raising=False)
    a = run(gpu, pts)
    monkeypatch.setenv("PARAGRAM_NUM_CHUNKS", "4")
    b = run(gpu, pts)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_c2_voronoi_1m_full_oracle(gpu):
    log("C2 voronoi 1M full oracle")
    pts = workloads.uniform(1_000_000, 1)
    got = run(gpu, pts)
    assert_matches_oracle(got, pts)


def test_c3_power_2m_full_oracle(gpu):
    log("C3 power 2M (clustered, hidden cells) full oracle")
    pts, w = workloads.clustered_power(2_000_000, 2)
    got = run(gpu, pts, w)
    csr_invariants(*got)
    assert (got[2] == 3).any(), "C3 is built to contain hidden cells"
    assert_matches_oracle(got, pts, w)


def test_c5_learned_power_10m_sample(gpu):
    log("C5 power 10M, 100k-cell seeded oracle sample")
    pts, w = workloads.learned_power(10_000_000, 4)
    adj, off, st = run(gpu, pts, w)
    sample = workloads.oracle_sample(pts.shape[0], 4, 100_000)
    ref = oracle.cells(pts, w, sample)
    bad = 0
    for k, i in enumerate(sample):
        bad += int(st[i] != ref.status[k] or not np.array_equal(adj[off[i]:off[i + 1]], ref.row(k)))
    assert bad == 0, f"{bad} of {sample.shape[0]} sampled cells differ"


def lattice_face_adjacency(m):
    """CSR of the face neighbours of an m x m x m cubic lattice (index = (x*m + y)*m + z)."""
    rows = []
    for x in range(m):
        for y in range(m):
            for z in range(m):
                nb = []
                for dx, dy, dz in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    u, v, w = x + dx, y + dy, z + dz
                    if 0 <= u < m and 0 <= v < m and 0 <= w < m:
                        nb.append((u * m + v) * m + w)
                rows.append(sorted(nb))
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return np.concatenate(rows).astype(np.int32), off


@pytest.mark.parametrize("weighted", [False, True])
def test_cubic_lattice_is_face_adjacency(gpu, weighted):
    """An exactly degenerate input (eight co-spherical points at every
    lattice vertex): every removal test of a diagonal neighbour is an exact
    tie.  The pre-clip offers the six face neighbours first, so no diagonal
    plane ever cuts and each row is exactly the cell's face neighbours.  (The
    rows of such inputs depend on the clipping order in the reference's
    semantics, which keep zero-area faces cut earlier; the oracle visits in
    another order, so this case is checked against the lattice itself.)"""
    log(f"cubic lattice 10^3, weighted={weighted}")
    m = 10
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) / 16.0
    w = np.full(g.shape[0], 1e-3, np.float32) if weighted else None  # equal weights: the same diagram
    adj, off, st = run(gpu, g, w)
    want_adj, want_off = lattice_face_adjacency(m)
    assert (st == 0).all()
    assert np.array_equal(off, want_off), "row lengths differ from the lattice's face counts"
    assert np.array_equal(adj, want_adj), "rows differ from the lattice's face neighbours"


def test_row_pool_growth_gives_identical_output(gpu):
    """A row pool far too small for the rows: the launch reports the size it
    needs and runs again; the output equals the default call's, and no cell
    is left truncated."""
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    log("row pool growth: 20k uniform and 20k power, pool of 1000 entries, unchunked and chunked")
    pts = workloads.uniform(20_000, 21)
    rng = np.random.default_rng(21)
    w = rng.uniform(0, 2e-4, 20_000).astype(np.float32)
    for weights in (None, w):
        p = torch.from_numpy(pts).to(gpu)
        wt = None if weights is None else torch.from_numpy(weights).reshape(-1, 1).to(gpu)
        want = [t.cpu().numpy() for t in _engine.build_diagram(_C.lib, p, wt)]
        for chunks in (None, "3"):
            if chunks:
                os.environ["PARAGRAM_NUM_CHUNKS"] = chunks
            try:
                ins = _engine.Instrument()
                got = [t.cpu().numpy() for t in _engine.build_diagram(_C.lib, p, wt, row_pool=1000, instrument=ins)]
            finally:
                os.environ.pop("PARAGRAM_NUM_CHUNKS", None)
            for x, y in zip(got, want):
                assert np.array_equal(x, y)
            assert ins.stats["pool_short"] == 0 and ins.stats["truncated"] == 0
            assert len(ins.cell_events) >= 2, "the small pool must have made a launch run again"


def test_long_equal_key_runs_cluster_and_outlier(gpu):
    """A dense cluster with one far outlier puts most points under a handful
    of Morton keys (runs of thousands): the duplicate search must stay linear
    and exact.  200k clustered points (with 2,000 exact duplicates) plus an
    outlier 1e4 cluster sizes away, against the oracle."""
    log("long equal-key runs: 200k cluster + outlier at 1e4 x the cluster size, with duplicates")
    rng = np.random.default_rng(31)
    c = (0.5 + 1e-4 * rng.standard_normal((200_000, 3))).astype(np.float32)
    c[-2000:] = c[:2000]  # exact duplicates inside the long runs
    pts = np.concatenate([c, np.array([[0.5 + 4.0, 0.5, 0.5]], np.float32)])
    got = run(gpu, pts)
    csr_invariants_fast(*got)
    assert_matches_oracle(got, pts)


def test_all_points_coincide(gpu):
    """Every point at the same position (zero extent: every Morton key equal)
    plus three distinct points, unweighted and with distinct weights."""
    log("coincident points: 5,000 copies of one point + 3 others")
    pts = np.concatenate([np.full((5_000, 3), 0.25, np.float32),
                          np.array([[0.5, 0.5, 0.5], [0.9, 0.1, 0.3], [0.2, 0.8, 0.7]], np.float32)])
    got = run(gpu, pts)
    assert_matches_oracle(got, pts)
    rng = np.random.default_rng(2)
    w = rng.uniform(0, 1e-3, pts.shape[0]).astype(np.float32)
    got = run(gpu, pts, w)
    assert_matches_oracle(got, pts, w)
