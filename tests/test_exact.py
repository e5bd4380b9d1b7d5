"""The library's exact removal predicate (expansion arithmetic, csrc/pg_exact.hpp).

CPU: the same code compiled for the host (pg_exact_side_host) against
rational arithmetic (Python Fractions) on random inputs, exact ties (where
the FP64 filter cannot decide) and their one-ulp neighbours, box planes and
wide exponent ranges.
GPU: the device build (pg_exact_side over an index) against the same
rationals and the CPU oracle's own expansion arithmetic.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from paragram_amd import workloads
from tests.helpers import exact_side_fractions, exact_tie_cloud


def _lib():
    from paragram_amd import _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    return _engine.HipLibrary()


def host_side(lib, points, weights, seed, tri, q, box=None):
    """pg_exact_side_host on one predicate given like exact_side_fractions."""
    pts = np.asarray(points, np.float32)
    w = np.zeros(len(pts), np.float32) if weights is None else np.asarray(weights, np.float32)
    lo, hi = oracle.box(pts) if box is None else box
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    seed4 = np.array([*pts[seed], w[seed]], np.float32)
    planes = np.zeros((4, 4), np.float32)
    codes = np.full(4, -1, np.int32)
    for e, o in enumerate(list(tri) + [q]):
        if o >= 0:
            planes[e] = [*pts[o], w[o]]
        else:
            codes[e] = -1 - o
    out = np.zeros(3, np.int32)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    lib.call("pg_exact_side_host", p(seed4), p(planes), p(codes), p(lo), p(hi), p(out))
    return tuple(int(x) for x in out)


def fp64_undecided(points, weights, seed, tri, q):
    """Does the FP64 filter of the cell kernels (|P| <= 1e-13 perm) leave this test open?"""
    pts = np.asarray(points, np.float64)
    w = np.zeros(len(pts)) if weights is None else np.asarray(weights, np.float64)
    s, ws = pts[seed], w[seed]
    planes = []
    for o in list(tri) + [q]:
        a = pts[o] - s
        planes.append((a, a @ a - (w[o] - ws), np.abs(a), a @ a + abs(w[o] - ws)))
    (a1, h1, p1, ph1), (a2, h2, p2, ph2), (a3, h3, p3, ph3), (aq, hq, pq, phq) = planes
    pc = lambda u, v: np.array([u[1] * v[2] + u[2] * v[1], u[2] * v[0] + u[0] * v[2], u[0] * v[1] + u[1] * v[0]])
    c1, c2, c3 = np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)
    q1, q2, q3 = pc(p2, p3), pc(p3, p1), pc(p1, p2)
    D, pD = a1 @ c1, p1 @ q1
    N, pN = h1 * c1 + h2 * c2 + h3 * c3, ph1 * q1 + ph2 * q2 + ph3 * q3
    P, pP = aq @ N - hq * D, pq @ pN + phq * pD
    return abs(P) <= 1e-13 * pP


def test_host_exact_matches_fractions_random():
    lib = _lib()
    rng = np.random.default_rng(3)
    pts = rng.random((40, 3)).astype(np.float32)
    w = rng.uniform(0, 1e-3, 40).astype(np.float32)
    for _ in range(200):
        idx = rng.choice(40, 5, replace=False)
        seed, tri, q = int(idx[0]), [int(x) for x in idx[1:4]], int(idx[4])
        assert host_side(lib, pts, w, seed, tri, q) == exact_side_fractions(pts, w, seed, tri, q)


def test_host_exact_matches_fractions_where_fp64_cannot_decide():
    """Exact ties (five points exactly co-spherical, or power-co-spherical
    with weights): P = 0, so the FP64 filter cannot decide, and one float ulp
    away from them.  (Float inputs rounded onto a sphere stay ~1e-8 off it,
    which FP64 decides; ties need exactly representable points.)"""
    lib = _lib()
    rng = np.random.default_rng(9)
    undecided = zeros = 0
    for trial in range(90):
        pts, w = exact_tie_cloud(rng, weighted=trial % 2 == 1)
        if trial % 3 == 0:
            pts[4] = np.nextafter(pts[4], np.float32(2.0))  # one ulp off the sphere
        undecided += int(fp64_undecided(pts, w, 0, [1, 2, 3], 4))
        for tri, q in (([1, 2, 3], 4), ([2, 1, 3], 4), ([1, 2, 4], 3)):
            want = exact_side_fractions(pts, w, 0, tri, q)
            zeros += int(want[2] == 0)
            assert host_side(lib, pts, w, 0, tri, q) == want
            assert oracle.exact_side(pts, w, 0, tri, q) == want
    assert undecided >= 30 and zeros >= 90, (undecided, zeros)


def test_host_exact_box_planes_and_weights():
    lib = _lib()
    pts = workloads.uniform(10, 1)
    w = np.linspace(0, 1e-2, 10).astype(np.float32)
    for weights in (None, w):
        for tri in ([-1, -3, -5], [-2, -4, 1], [-1, 2, 3], [-6, -3, 7]):
            for q in (4, 5, 6, -2, -5):
                if q in tri:
                    continue
                assert host_side(lib, pts, weights, 0, tri, q) == exact_side_fractions(pts, weights, 0, tri, q)


def test_host_exact_wide_exponent_range():
    """Coordinates and weights spread over many binades: differences of the
    FP32 inputs need two doubles, and the products many components."""
    lib = _lib()
    rng = np.random.default_rng(17)
    for _ in range(150):
        scale = 10.0 ** rng.integers(-30, 6, size=(6, 1))
        pts = (0.5 + rng.standard_normal((6, 3)) * scale).astype(np.float32)
        pts[0] = np.float32(0.5)
        w = (rng.random(6) * 10.0 ** rng.integers(-38, 0, size=6)).astype(np.float32)
        tri, q = [1, 2, 3], 4 + int(rng.integers(0, 2))
        got = host_side(lib, pts, w, 0, tri, q)
        assert got == exact_side_fractions(pts, w, 0, tri, q)


def test_planted_sign_mistake_is_caught():
    """A mistake planted in the rational reference (the opposite orientation of
    one vertex) is seen by the same comparison the tests above use."""
    lib = _lib()
    rng = np.random.default_rng(4)
    pts = rng.random((12, 3)).astype(np.float32)
    want = exact_side_fractions(pts, None, 0, [1, 2, 3], 4)
    planted = exact_side_fractions(pts, None, 0, [2, 1, 3], 4)  # the swapped triple flips sign(D)
    got = host_side(lib, pts, None, 0, [1, 2, 3], 4)
    assert got == want and got != planted


@pytest.mark.gpu
def test_device_exact_side_matches_fractions_and_oracle(gpu):
    """pg_exact_side (the device build of the same code) over an index of a
    cloud, on predicates FP64 cannot decide and on random ones."""
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    lib = _C.lib
    rng = np.random.default_rng(21)
    clouds, queries = [], []
    for trial in range(40):
        tie, _ = exact_tie_cloud(rng)
        if trial % 2 == 0:
            tie[4] = np.nextafter(tie[4], np.float32(2.0))
        clouds.append(np.concatenate([tie, rng.random((3, 3)).astype(np.float32)]))
    # one index per cloud (box from the cloud), all queries of that cloud in one launch
    for pts in clouds:
        qs = [(0, 1, 2, 3, 4), (0, 2, 1, 3, 4), (0, 1, 2, 4, 3), (5, 6, 7, 1, 2), (0, -1, -3, -5, 6),
              (1, -2, 3, 4, 0)]
        dev = torch.device(gpu)
        p = torch.from_numpy(pts).to(dev)
        n = pts.shape[0]
        ws_bytes = lib.u64("pg_index_workspace_bytes", n)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        lib.call("pg_build_index", n, _engine._p(p), None, -1, _engine._p(ws), ws_bytes, sp)
        qt = torch.tensor(qs, dtype=torch.int32, device=dev)
        out = torch.empty((len(qs), 3), dtype=torch.int32, device=dev)
        lib.call("pg_exact_side", n, _engine._p(ws), len(qs), _engine._p(qt), _engine._p(out), sp)
        got = out.cpu().numpy()
        for k, (s, a, b, c, q) in enumerate(qs):
            want = exact_side_fractions(pts, None, s, [a, b, c], q)
            assert tuple(int(x) for x in got[k]) == want, (k, qs[k])
            assert oracle.exact_side(pts, None, s, [a, b, c], q) == want
        queries.append(len(qs))
    assert sum(queries) == 240
