"""GPU kNN against brute force: the k nearest neighbours of every point, the
point itself excluded, ascending squared distance, short lists padded with
the last id (or 0), as the reference's kNN pass (pwr_bvh.cu:159-215)."""
import numpy as np
import pytest

from paragram_amd import workloads

pytestmark = pytest.mark.gpu


def run_knn(gpu, pts, k):
    import torch

    from paragram_amd import _backend, _engine

    ids, d2 = _engine.knn(_backend._C.lib, torch.from_numpy(np.ascontiguousarray(pts)).to(gpu), k)
    return ids.cpu().numpy(), d2.cpu().numpy()


def brute(pts, k):
    p = pts.astype(np.float64)
    n = p.shape[0]
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    d2[np.arange(n), np.arange(n)] = np.inf
    order = np.argsort(d2, axis=1, kind="stable")[:, :k]
    return order, np.take_along_axis(d2, order, 1)


@pytest.mark.parametrize("k", [8, 16, 3])
def test_knn_uniform_matches_brute_force(gpu, k):
    print(f"knn uniform 3000 k={k}", flush=True)
    pts = workloads.uniform(3000, 7)
    ids, d2 = run_knn(gpu, pts, k)
    bid, bd2 = brute(pts, k)
    assert np.allclose(d2, bd2, rtol=2e-6, atol=0), "k nearest squared distances differ"
    # ids agree wherever the k-th and (k+1)-th distances are not tied
    p = pts.astype(np.float64)
    got = ((p[ids] - p[:, None, :]) ** 2).sum(-1)
    assert np.allclose(got, bd2, rtol=2e-6), "returned ids are not at the returned distances"
    assert (ids != np.arange(pts.shape[0])[:, None]).all(), "a point is its own neighbour"


def test_knn_lidar_cloud(gpu):
    print("knn on a LiDAR-like cloud (deep hierarchy, dense rings)", flush=True)
    cloud = workloads.lidar_batch(3, max_clouds=1)[0][:20000]
    ids, d2 = run_knn(gpu, cloud, 8)
    sub = np.random.default_rng(0).choice(cloud.shape[0], 300, replace=False)
    p = cloud.astype(np.float64)
    for i in sub:
        dd = ((p - p[i]) ** 2).sum(-1)
        dd[i] = np.inf
        want = np.sort(dd)[:8]
        assert np.allclose(d2[i], want, rtol=2e-6), f"point {i}"


@pytest.mark.parametrize("n", [1, 2, 5, 9])
def test_knn_short_lists_are_padded(gpu, n):
    print(f"knn padding n={n}", flush=True)
    pts = workloads.uniform(n, 40 + n)
    ids, d2 = run_knn(gpu, pts, 8)
    for i in range(n):
        found = n - 1
        if found == 0:
            assert (ids[i] == 0).all()
            continue
        assert (ids[i, found:] == ids[i, found - 1]).all(), "padding repeats the last id"
        assert (d2[i, found:] > 1e30).all()
        assert sorted(ids[i, :found]) == sorted(set(range(n)) - {i})


def test_knn_duplicates_are_neighbours(gpu):
    print("knn with duplicate points", flush=True)
    pts = workloads.uniform(500, 9)
    pts = np.concatenate([pts, pts[:50]])
    ids, d2 = run_knn(gpu, pts, 8)
    for i in range(50):
        assert d2[i, 0] == 0.0 and ids[i, 0] == 500 + i
        assert d2[500 + i, 0] == 0.0 and ids[500 + i, 0] == i


def chain_cloud():
    """A hierarchy deeper than the first pass's 32-entry stack: 4096
    coincident points at the origin (split by index below the 30 Morton bit
    levels) and one point on each Morton bit level (x, y or z = 2^-k), so the
    path to the origin cluster is about 37 levels, and a first descent (every
    far child still within the unset k-th distance) pushes at each."""
    pts = [np.zeros((4096, 3))]
    for k in range(1, 11):
        for axis in range(3):
            p = np.zeros(3)
            p[axis] = 2.0 ** -k
            pts.append(p[None])
    pts.append(np.ones((1, 3)))
    return np.concatenate(pts).astype(np.float32)


def test_knn_deep_hierarchy_spills_to_the_deep_pass(gpu):
    import torch

    from paragram_amd import _backend, _engine

    print("knn on a 37-level hierarchy (first-pass stack overflow -> deep pass)", flush=True)
    pts = chain_cloud()
    n, k = pts.shape[0], 16
    lib = _backend._C.lib
    P = torch.from_numpy(pts).to(gpu)
    ws = torch.empty(lib.u64("pg_index_workspace_bytes", n), dtype=torch.uint8, device=gpu)
    ids = torch.empty((n, k), dtype=torch.int32, device=gpu)
    d2 = torch.empty((n, k), dtype=torch.float32, device=gpu)
    import ctypes

    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib.call("pg_build_index", n, _engine._p(P), None, -1, _engine._p(ws), ws.numel(), sp)
    lib.call("pg_knn", n, _engine._p(ws), k, _engine._p(ids), _engine._p(d2), sp)
    torch.cuda.synchronize()
    # IndexHeader: 12 floats, then root_ref, n, has_weights, leaf_size,
    # has_duplicates, knn_spill (uint32 each): knn_spill at byte 68
    spilled = int(ws[68:72].cpu().numpy().view(np.uint32)[0])
    print(f"queries redone by the deep pass: {spilled}", flush=True)
    assert spilled > 0, "the dataset did not reach the deep pass"
    ids, d2 = ids.cpu().numpy(), d2.cpu().numpy()
    _, bd2 = brute(pts, k)
    assert np.allclose(d2, bd2, rtol=2e-6, atol=0), "k nearest squared distances differ"
    p = pts.astype(np.float64)
    got = ((p[ids] - p[:, None, :]) ** 2).sum(-1)
    assert np.allclose(got, bd2, rtol=2e-6, atol=0), "returned ids are not at the returned distances"
    assert (ids != np.arange(n)[:, None]).all(), "a point is its own neighbour"
