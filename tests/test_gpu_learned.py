"""GPU checks of the learned-radius path: the matrix-core forward and the
training kernels against the float64 reference (tests/mlp_ref.py, itself
pinned by finite differences), the features against numpy, and the learned
build against the plain build (identical diagrams, less work)."""
import numpy as np
import pytest

from paragram_amd import workloads
from tests import mlp_ref

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def lib():
    from paragram_amd import _backend

    return _backend._C.lib


def rand_params(seed):
    from paragram_amd import learned

    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    o = 0
    for fi, fo in mlp_ref.SIZES:  # non-zero biases exercise the bias paths
        o += fi * fo
        p[o:o + fo] = rng.normal(0, 0.2, fo)
        o += fo
    return p.astype(np.float32)


def test_mlp_forward_matches_reference(gpu):
    import torch

    from paragram_amd import learned

    log("mlp forward (MFMA) vs float64 reference, 1000 rows")
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (1000, 16)).astype(np.float32)
    p = rand_params(4)
    Y = learned.forward(lib(), torch.from_numpy(X).to(gpu), torch.from_numpy(p).to(gpu)).cpu().numpy()
    ref, _ = mlp_ref.forward(p, X, bf16_ops=True)
    err = np.abs(Y - ref)
    assert np.median(err) < 1e-4 * (1 + np.median(np.abs(ref))), np.median(err)
    assert (err <= 2e-2 * (1 + np.abs(ref))).all(), err.max()


def test_mlp_gradient_and_adam_match_reference(gpu):
    import torch

    from paragram_amd import learned

    log("mlp training step (gradient + Adam) vs float64 reference")
    rng = np.random.default_rng(5)
    n = 700
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    T = rng.normal(0, 1, (n, 2)).astype(np.float32)
    mask = (rng.random(n) > 0.1).astype(np.float32)
    p = rand_params(6)
    t = learned.Trainer(lib(), p, gpu, n)
    Xd, Td, Md = (torch.from_numpy(a).to(gpu) for a in (X, T, mask))
    loss = float(t.step(Xd, Td, Md, float(mask.sum()), 0.0))
    g = t.grad[:5378].cpu().numpy().astype(np.float64)
    ref_loss, ref_g = mlp_ref.loss_and_grad(p, X, T, mask, bf16_ops=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    assert np.linalg.norm(g - ref_g) <= 1e-2 * np.linalg.norm(ref_g)
    # one Adam step from the same point
    t2 = learned.Trainer(lib(), p, gpu, n)
    t2.step(Xd, Td, Md, float(mask.sum()), 1e-2)
    got = t2.P.cpu().numpy().astype(np.float64)
    want, _, _ = mlp_ref.adam(p.astype(np.float64), t2.grad[:5378].cpu().numpy().astype(np.float64),
                              np.zeros(5378), np.zeros(5378), 1, 1e-2)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)


def test_features_match_numpy(gpu):
    import torch

    from paragram_amd import learned

    log("features vs numpy (power cloud, 3000 points)")
    pts, w = workloads.clustered_power(3000, 8)
    X, logd8 = learned.features(lib(), torch.from_numpy(pts).to(gpu), torch.from_numpy(w).to(gpu).reshape(-1, 1))
    X, logd8 = X.cpu().numpy(), logd8.cpu().numpy()
    p = pts.astype(np.float64)
    lo, hi = p.min(0), p.max(0)
    for i in np.random.default_rng(0).choice(3000, 200, replace=False):
        d2 = ((p - p[i]) ** 2).sum(1)
        d2[i] = np.inf
        nn = np.argsort(d2, kind="stable")[:8]
        d = np.sqrt(d2[nn])
        d8 = d[7]
        wn = w[nn].astype(np.float64)
        b = max(min((p[i] - lo).min(), (hi - p[i]).min()), 0) / d8
        want = np.concatenate([d[:7] / d8, [np.linalg.norm((p[nn] - p[i]).mean(0)) / d8, np.log1p(b),
                               min(1.0, b / 8), (w[i] - wn.mean()) / d8 ** 2, (wn.max() - w[i]) / d8 ** 2,
                               (w[i] - wn.min()) / d8 ** 2, wn.std() / d8 ** 2, (d8 - d[0]) / d8, 1.0]])
        assert np.allclose(X[i], want, rtol=1e-4, atol=1e-4), (i, X[i], want)
        assert abs(logd8[i] - np.log(d8)) < 1e-5


def run_plain_and_cube(gpu, pts, w, h, knn=8):
    import torch

    from paragram_amd import _engine

    P = torch.from_numpy(pts).to(gpu)
    W = None if w is None else torch.from_numpy(w).to(gpu).reshape(-1, 1)
    a, b = _engine.Instrument(), _engine.Instrument()
    plain = _engine.build_diagram(lib(), P, W, instrument=a, knn=knn)
    cube = _engine.build_diagram(lib(), P, W, instrument=b, init_h=h, knn=knn)
    return plain, cube, a.stats, b.stats


@pytest.mark.parametrize("mode", ["tiny", "random"])
def test_wrong_predictions_do_not_change_the_diagram(gpu, mode):
    import torch

    log(f"cube start with {mode} predictions equals the plain diagram")
    pts, w = workloads.learned_power(200_000, 31)
    rng = np.random.default_rng(2)
    h = (np.full(200_000, 1e-4) if mode == "tiny" else rng.uniform(0, 0.05, 200_000)).astype(np.float32)
    plain, cube, _, _ = run_plain_and_cube(gpu, pts, w, torch.from_numpy(h).to(gpu))
    for x, y in zip(plain, cube):
        assert torch.equal(x, y)


def test_trained_network_gain(gpu):
    import torch

    from paragram_amd import learned

    log("train a small network on 2 clouds, then learned vs plain on a third")
    clouds = [workloads.learned_power(300_000, 40 + k) for k in range(2)]
    model = learned.train(lib(), clouds, gpu, epochs=4, batch=2048, rows_per_cloud=300_000, log=log)
    pts, w = workloads.learned_power(500_000, 50)
    P, W = torch.from_numpy(pts).to(gpu), torch.from_numpy(w).to(gpu).reshape(-1, 1)
    h = learned.predict_half_width(lib(), model, P, W)
    plain, cube, _, _ = run_plain_and_cube(gpu, pts, w, h)
    for x, y in zip(plain, cube):
        assert torch.equal(x, y), "the learned start changed the diagram"
    # the gain is measured without the kNN pre-clip, where the predicted cube
    # is the only starting bound (on top of the pre-clip it saves nothing)
    bare, cube, sp, sc = run_plain_and_cube(gpu, pts, w, h, knn=0)
    for x, y in zip(bare, cube):
        assert torch.equal(x, y), "the learned start changed the diagram"
    work_p = sp["clips"] + sp["candidates"]
    work_c = sc["clips"] + sc["candidates"]
    log(f"plain: clips {sp['clips']} candidates {sp['candidates']}; learned: clips {sc['clips']} "
        f"candidates {sc['candidates']} ({100 * (1 - work_c / work_p):.1f}% less clip work)")
    assert work_c < work_p
    d = learned.power_diagram(P, W, model=model)
    assert torch.equal(d.adjacency, plain[0]) and torch.equal(d.offsets, plain[1])


def test_c5_learned_form_with_committed_model(gpu):
    """The learned-radius power configuration (10M points, seed 4) with the
    committed network (paragram_amd/radius_mlp.npz): byte-identical to the
    plain call, and a seeded 100k-cell sample equal to the oracle."""
    import torch

    import oracle
    import paragram_amd
    from paragram_amd import learned

    log("C5 10M: learned form (committed radius_mlp.npz) vs plain vs 100k-cell oracle sample")
    pts, w = workloads.learned_power(10_000_000, 4)
    P, W = torch.from_numpy(pts).to(gpu), torch.from_numpy(w).to(gpu)
    model = learned.load()
    d = learned.power_diagram(P, W, model=model)
    got = (d.adjacency.cpu().numpy(), d.offsets.cpu().numpy(), d.status.cpu().numpy())
    del d
    plain = paragram_amd.power_diagram(P, W)
    for name, x, y in zip(("adjacency", "offsets", "status"), got, (plain.adjacency, plain.offsets, plain.status)):
        assert np.array_equal(x, y.cpu().numpy()), f"learned form: {name} differs from the plain call"
    del plain
    adj, off, st = got
    sample = workloads.oracle_sample(pts.shape[0], 4, 100_000)
    ref = oracle.cells(pts, w, sample)
    bad = sum(int(st[i] != ref.status[k] or not np.array_equal(adj[off[i]:off[i + 1]], ref.row(k)))
              for k, i in enumerate(sample))
    assert bad == 0, f"{bad} of {sample.shape[0]} sampled cells differ"


def test_learned_batched_form_equals_plain_batch(gpu):
    """The learned radius on batches (neighbours and hull features per cloud):
    32 LiDAR clouds (seed 3) and a ragged batch of power clouds give exactly
    the plain batched results, cloud by cloud, and the prediction on the
    batch index equals each cloud's prediction on its own index."""
    import ctypes

    import torch

    import paragram_amd
    from paragram_amd import _engine, learned

    log("learned radius on batches: 32 LiDAR clouds and 6 power clouds vs the plain batched calls")
    model = learned.load()
    cl = workloads.lidar_batch(3, n_clouds=256, max_clouds=32)
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in cl])]).astype(np.int64)
    P = torch.from_numpy(np.concatenate(cl)).to(gpu)
    got = learned.batched_voronoi_diagram(P, off, model=model)
    want = paragram_amd.batched_voronoi_diagram(P, off)
    assert len(got) == len(want) == 32
    for c, (g, w) in enumerate(zip(got, want)):
        for name in ("adjacency", "offsets", "status"):
            assert torch.equal(getattr(g, name), getattr(w, name)), (c, name)
    rng = np.random.default_rng(11)
    sizes = [3000, 1, 0, 5000, 17, 8000]
    pts, ws = [], []
    for s in sizes:
        pw = workloads.learned_power(max(s, 1), int(rng.integers(1 << 30)))
        pts.append(pw[0][:s])
        ws.append(pw[1][:s])
    off2 = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    P2 = torch.from_numpy(np.concatenate(pts)).to(gpu)
    W2 = torch.from_numpy(np.concatenate(ws)).to(gpu)
    got = learned.batched_power_diagram(P2, W2, off2, model=model)
    want = paragram_amd.batched_power_diagram(P2, W2, off2)
    for c, (g, w) in enumerate(zip(got, want)):
        for name in ("adjacency", "offsets", "status"):
            assert torch.equal(getattr(g, name), getattr(w, name)), (c, name)
    # the batch prediction itself (per-cloud kNN and hull features on the batch
    # index) equals each cloud's prediction on its own index
    lib = learned._lib()
    sp = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    n = P2.shape[0]
    ws_bytes = lib.u64("pg_batch_workspace_bytes", n, len(sizes))
    wsb = torch.empty(ws_bytes, dtype=torch.uint8, device=gpu)
    lib.call("pg_build_batch_index", n, len(sizes), _engine._p(P2), _engine._p(W2), ctypes.c_void_p(off2.ctypes.data),
             -1, _engine._p(wsb), ws_bytes, sp)
    hb = learned.predict_on_index(lib, model, wsb, n, gpu, sp, len(sizes)).cpu().numpy()
    for c, sz in enumerate(sizes):
        if sz < 17:
            continue
        a = int(off2[c])
        Pc, Wc = P2[a:a + sz].contiguous(), W2[a:a + sz].contiguous()
        wb = lib.u64("pg_index_workspace_bytes", sz)
        wsc = torch.empty(wb, dtype=torch.uint8, device=gpu)
        lib.call("pg_build_index", sz, _engine._p(Pc), _engine._p(Wc), -1, _engine._p(wsc), wb, sp)
        hc = learned.predict_on_index(lib, model, wsc, sz, gpu, sp).cpu().numpy()
        assert np.allclose(hb[a:a + sz], hc, rtol=1e-5, atol=0), c
