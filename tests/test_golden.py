"""The committed golden fixtures (tests/golden/, written and cross-checked by
tests/golden/make_golden.py): the oracle and the library's host exact-sign
code reproduce them on the CPU; the HIP path reproduces them on the GPU."""
import ctypes
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIAGRAMS = ["voronoi_uniform_300", "power_hidden_400", "power_clustered_500"]


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", DIAGRAMS)
def test_oracle_reproduces_golden_diagram(name):
    g = load(name)
    adj, off, st, _ = oracle.diagram(g["points"], g.get("weights"))
    assert np.array_equal(adj, g["adjacency"]) and np.array_equal(off, g["offsets"])
    assert np.array_equal(st, g["status"])


def predicate_rows():
    g = load("predicates")
    return g["points"], g["weights"], g["expected"]


def test_oracle_and_host_exact_sign_reproduce_golden_predicates():
    from paragram_amd import _engine
    from tests.test_exact import host_side

    P, W, E = predicate_rows()
    lib = _engine.HipLibrary() if os.path.exists(_engine.LIB_PATH) else None
    for k in range(P.shape[0]):
        want = tuple(int(x) for x in E[k])
        assert oracle.exact_side(P[k], W[k], 0, [1, 2, 3], 4) == want, k
        if lib is not None:
            assert host_side(lib, P[k], W[k], 0, [1, 2, 3], 4) == want, k


@pytest.mark.gpu
@pytest.mark.parametrize("name", DIAGRAMS + ["lattice_4"])
def test_hip_path_reproduces_golden_diagram(gpu, name):
    import torch

    import paragram_amd

    g = load(name)
    p = torch.from_numpy(g["points"]).to(gpu)
    if "weights" in g:
        d = paragram_amd.power_diagram(p, torch.from_numpy(g["weights"]).to(gpu))
    else:
        d = paragram_amd.voronoi_diagram(p)
    assert np.array_equal(d.adjacency.cpu().numpy(), g["adjacency"])
    assert np.array_equal(d.offsets.cpu().numpy(), g["offsets"])
    assert np.array_equal(d.status.cpu().numpy(), g["status"])


@pytest.mark.gpu
def test_device_exact_sign_reproduces_golden_predicates(gpu):
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    lib = _C.lib
    P, W, E = predicate_rows()
    sp = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    q = torch.tensor([[0, 1, 2, 3, 4]], dtype=torch.int32, device=gpu)
    out = torch.empty((1, 3), dtype=torch.int32, device=gpu)
    for k in range(P.shape[0]):
        pts = torch.from_numpy(P[k]).to(gpu)
        w = torch.from_numpy(W[k]).to(gpu)
        ws_bytes = lib.u64("pg_index_workspace_bytes", 5)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=gpu)
        lib.call("pg_build_index", 5, _engine._p(pts), _engine._p(w), -1, _engine._p(ws), ws_bytes, sp)
        lib.call("pg_exact_side", 5, _engine._p(ws), 1, _engine._p(q), _engine._p(out), sp)
        assert tuple(int(x) for x in out[0].cpu()) == tuple(int(x) for x in E[k]), k
