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
from fractions import Fraction

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


def host_vertex(lib, seed4, planes, codes, lo, hi):
    """pg_exact_vertex_host: (y0, y1, y2, w) of the vertex of three planes."""
    out = np.zeros(4, np.float32)
    arrs = [np.ascontiguousarray(a, dt) for a, dt in ((seed4, np.float32), (planes, np.float32),
                                                       (codes, np.int32), (lo, np.float32), (hi, np.float32))]
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    lib.call("pg_exact_vertex_host", *(p(a) for a in arrs), p(out))
    return out


def vertex_fractions(seed4, planes, codes, lo, hi):
    """(D, y) of the vertex of three planes in rational arithmetic (y None when D == 0)."""
    s = [Fraction(float(v)) for v in seed4]
    A, H = [], []
    for pl, code in zip(planes, codes):
        if code < 0:
            a = [Fraction(float(pl[k])) - s[k] for k in range(3)]
            A.append(a)
            H.append(sum(x * x for x in a) - (Fraction(float(pl[3])) - s[3]))
        else:
            axis, side = int(code) // 2, int(code) % 2
            a = [Fraction(0)] * 3
            a[axis] = Fraction(1 if side else -1)
            A.append(a)
            H.append(2 * (Fraction(float(hi[axis])) - s[axis]) if side else 2 * (s[axis] - Fraction(float(lo[axis]))))
    cr = lambda u, v: [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    c = [cr(A[1], A[2]), cr(A[2], A[0]), cr(A[0], A[1])]
    D = sum(A[0][k] * c[0][k] for k in range(3))
    if D == 0:
        return D, None
    return D, [sum(H[t] * c[t][k] for t in range(3)) / D for k in range(3)]


def test_host_exact_vertex_position_matches_fractions():
    """The exact vertex position of the FP64-capable tiers: within its stated
    radius of the rational vertex, that radius within FP32 rounding, and no
    vertex (-1) for a non-positive orientation.  Cases: nearly parallel
    planes of three close points far from the seed (the LiDAR ground cell
    whose FP32 radius was 25 at |y| = 5), nearly coplanar points, box planes,
    weights, and random triples."""
    lib = _lib()
    h = float.fromhex
    lidar_seed = np.array([h("0x1.a564bp+3"), h("0x1.d280d8p+4"), h("-0x1.cd0d2cp+0"), 0.0], np.float32)
    lidar_pts = np.array([[h("0x1.ef3bd8p+3"), h("0x1.0d9fc6p+5"), h("-0x1.cca5d4p+0"), 0.0],
                          [h("0x1.eaf82cp+3"), h("0x1.0e1dc8p+5"), h("-0x1.ccaaf2p+0"), 0.0],
                          [h("0x1.f367ccp+3"), h("0x1.0d1118p+5"), h("-0x1.cc955ep+0"), 0.0]], np.float32)
    lo = np.array([-81.0, -81.0, -2.83], np.float32)
    hi = np.array([81.0, 81.0, 5.19], np.float32)
    cases = [(lidar_seed, lidar_pts, np.full(3, -1, np.int32))]
    rng = np.random.default_rng(33)
    for trial in range(300):
        seed = np.array([*rng.uniform(-1, 1, 3), 0.0], np.float32)
        kind = trial % 4
        if kind == 0:  # three close points in one direction: nearly parallel planes
            base = seed[:3] + rng.normal(0, 1, 3) * 3
            pts = base + rng.normal(0, 1, (3, 3)) * 10.0 ** rng.uniform(-4, -1)
        elif kind == 1:  # nearly coplanar with the seed: nearly coaxial planes
            u, v = rng.normal(0, 1, 3), rng.normal(0, 1, 3)
            coef = rng.normal(0, 1, (3, 2))
            pts = seed[:3] + coef[:, :1] * u + coef[:, 1:] * v + rng.normal(0, 1, (3, 3)) * 1e-5
        else:
            pts = seed[:3] + rng.normal(0, 1, (3, 3))
        planes = np.zeros((3, 4), np.float32)
        planes[:, :3] = pts
        if kind == 3:
            planes[:, 3] = rng.uniform(0, 0.1, 3)
            seed[3] = np.float32(0.05)
        codes = np.full(3, -1, np.int32)
        if kind == 2 and trial % 8 == 2:
            codes[rng.integers(0, 3)] = rng.integers(0, 6)
        cases.append((seed, planes, codes))
    box = (np.array([-40, -40, -40], np.float32), np.array([40, 40, 40], np.float32))
    seen_pos = seen_neg = 0
    for k, (seed, planes, codes) in enumerate(cases):
        blo, bhi = (lo, hi) if k == 0 else box
        for order in ((0, 1, 2), (0, 2, 1)):
            pl, cd = planes[list(order)], codes[list(order)]
            out = host_vertex(lib, seed, pl, cd, blo, bhi)
            D, y = vertex_fractions(seed, pl, cd, blo, bhi)
            if D <= 0:
                assert out[3] == -1.0, (k, order)
                seen_neg += 1
                continue
            seen_pos += 1
            w = Fraction(float(out[3]))
            ym = max(abs(v) for v in y)
            for j in range(3):
                assert abs(Fraction(float(out[j])) - y[j]) <= w, (k, order, j)
            # FP32 rounding of the position plus the 1e-13 term, nothing more
            assert w <= ym * Fraction(1, 2**23) + Fraction(1, 10**20), (k, order, float(w), float(ym))
    assert seen_pos >= 290 and seen_neg >= 290
    out = host_vertex(lib, lidar_seed, lidar_pts, np.full(3, -1, np.int32), lo, hi)
    assert out[3] < 1e-6 and abs(out[2] - (-1.9751966)) < 1e-6


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
