"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no tests or golden vectors, so the
oracle is checked against independent constructions:

* Qhull (scipy): the Voronoi adjacency of points clipped to a box equals the
  Delaunay adjacency of the points together with their mirror images across
  the six box planes (a mirror never beats its original inside the box, and a
  point's own mirrors contribute exactly the box planes).  The power diagram
  is the lower convex hull of the lifted points (x, |x|^2 - w) with the same
  mirrors; hidden points are those missing from the lower hull.
* Brute force: every triple of planes of a cell (bisectors of all points and
  the box planes), its intersection point, kept when inside every half-space;
  a neighbour is an owner with at least three distinct vertices on its plane.
* Exact arithmetic: the removal predicate sign against Python Fractions on
  the same float inputs, including near-degenerate configurations that the
  FP64 filter cannot decide.
* The FP32 box of voronoi_ultra.cu:237-238 / voronoi_convex_cell.cuh:187-193.

A test that plants a mistake does so in a copy of the oracle's output (a
dropped neighbour, a flipped status) and checks that the comparison catches it.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay

import oracle
from paragram_amd import workloads
from tests.helpers import exact_side_fractions as _exact_side
from tests.helpers import csr_invariants_fast, exact_tie_cloud
from tests.helpers import csr_invariants, rows_of


def qhull_rows(points, weights=None):
    pts = np.asarray(points, np.float64)
    n = len(pts)
    lo, hi = oracle.box(points)
    w = np.zeros(n) if weights is None else np.asarray(weights, np.float64)
    allp, allw = [pts], [w]
    for ax in range(3):
        for b in (float(lo[ax]), float(hi[ax])):
            m = pts.copy()
            m[:, ax] = 2 * b - m[:, ax]
            allp.append(m)
            allw.append(w)
    P = np.concatenate(allp)
    W = np.concatenate(allw)
    if weights is None:
        simp = Delaunay(P).simplices
    else:
        lifted = np.concatenate([P, (P * P).sum(1, keepdims=True) - W[:, None]], 1)
        hull = ConvexHull(lifted)
        simp = hull.simplices[hull.equations[:, 3] < 0]
    rows = [set() for _ in range(n)]
    for sx in simp:
        real = [int(a) for a in sx if a < n]
        for a in real:
            rows[a].update(b for b in real if b != a)
    present = np.zeros(n, bool)
    present[np.unique(simp[simp < n])] = True
    return [np.array(sorted(r), np.int32) for r in rows], present


def brute_rows(points, weights=None):
    """O(N^4)-per-cell brute-force clipping in float64 (tiny N only)."""
    pts = np.asarray(points, np.float64)
    n = len(pts)
    w = np.zeros(n) if weights is None else np.asarray(weights, np.float64)
    lo, hi = oracle.box(points)
    rows, status = [], []
    for s in range(n):
        A, H, own = [], [], []
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = -1.0
            A.append(e.copy()); H.append(2 * (pts[s, ax] - float(lo[ax]))); own.append(-1)
            e[ax] = 1.0
            A.append(e.copy()); H.append(2 * (float(hi[ax]) - pts[s, ax])); own.append(-1)
        for p in range(n):
            if p == s:
                continue
            a = pts[p] - pts[s]
            A.append(a); H.append(a @ a - (w[p] - w[s])); own.append(p)
        A = np.array(A); H = np.array(H)
        m = len(A)
        verts = []
        for i in range(m):
            for j in range(i + 1, m):
                for k in range(j + 1, m):
                    M = A[[i, j, k]]
                    if abs(np.linalg.det(M)) < 1e-12:
                        continue
                    y = np.linalg.solve(M, H[[i, j, k]])
                    if np.all(A @ y <= H + 1e-9 * (1 + np.abs(H))):
                        verts.append(y)
        if not verts:
            rows.append(np.zeros(0, np.int32)); status.append(3)
            continue
        V = np.array(verts)
        nb = []
        for t in range(m):
            if own[t] < 0:
                continue
            on = np.abs(V @ A[t] - H[t]) <= 1e-9 * (1 + abs(H[t]))
            if len(np.unique(np.round(V[on], 9), axis=0)) >= 3:
                nb.append(own[t])
        rows.append(np.array(sorted(nb), np.int32)); status.append(0)
    return rows, np.array(status, np.int32)


def test_box_matches_fp32_reference_order():
    pts = workloads.uniform(100, 3)
    lo, hi = oracle.box(pts)
    mn = pts.min(0)
    mx = pts.max(0)
    exp_lo = ((mn - np.float32(1.0)).astype(np.float32) - np.float32(1e-5)).astype(np.float32)
    exp_hi = ((mx + np.float32(1.0)).astype(np.float32) + np.float32(1e-5)).astype(np.float32)
    assert np.array_equal(lo, exp_lo) and np.array_equal(hi, exp_hi)


@pytest.mark.parametrize("n,seed", [(300, 0), (1500, 1), (4000, 2)])
def test_voronoi_matches_qhull(n, seed):
    pts = workloads.uniform(n, seed)
    r = oracle.cells(pts)
    q, _ = qhull_rows(pts)
    for i in range(n):
        assert np.array_equal(r.row(i), q[i]), f"cell {i}"
    assert np.all(r.status == 0)


@pytest.mark.parametrize("n,seed,wmax", [(400, 5, 0.02), (1500, 6, 0.01), (3000, 7, 4e-3)])
def test_power_matches_lifted_hull(n, seed, wmax):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3)).astype(np.float32)
    w = rng.uniform(0, wmax, n).astype(np.float32)
    r = oracle.cells(pts, w)
    q, present = qhull_rows(pts, w)
    assert np.array_equal(r.status == 3, ~present)
    assert (r.status == 3).sum() > 0
    for i in range(n):
        assert np.array_equal(r.row(i), q[i]), f"cell {i}"


def test_clustered_clamped_power_matches_lifted_hull():
    pts, w = workloads.clustered_power(3000, 2)
    r = oracle.cells(pts, w)
    q, present = qhull_rows(pts, w)
    assert np.array_equal(r.status == 3, ~present)
    for i in range(len(pts)):
        assert np.array_equal(r.row(i), q[i])


def test_lidar_like_cloud_matches_qhull():
    cloud = workloads.lidar_batch(3, max_clouds=1)[0][:3000]
    r = oracle.cells(cloud)
    q, _ = qhull_rows(cloud)
    for i in range(len(cloud)):
        assert np.array_equal(r.row(i), q[i])


@pytest.mark.parametrize("weighted", [False, True])
def test_brute_force_tiny(weighted):
    rng = np.random.default_rng(11 + weighted)
    pts = rng.random((14, 3)).astype(np.float32)
    w = rng.uniform(0, 0.05, 14).astype(np.float32) if weighted else None
    r = oracle.cells(pts, w)
    rows, st = brute_rows(pts, w)
    assert np.array_equal(r.status, st)
    for i in range(len(pts)):
        assert np.array_equal(r.row(i), rows[i]), f"cell {i}"


def test_single_point_and_pair():
    one = np.array([[0.5, 0.5, 0.5]], np.float32)
    r = oracle.cells(one)
    assert r.counts.tolist() == [0] and r.status.tolist() == [0]
    two = np.array([[0.2, 0.5, 0.5], [0.8, 0.5, 0.5]], np.float32)
    adj, off, st, _ = oracle.diagram(two)
    assert adj.tolist() == [1, 0] and off.tolist() == [0, 1, 2] and st.tolist() == [0, 0]


def test_hidden_point_power():
    # the middle point's weight is so small that its neighbours' cells cover it
    pts = np.array([[0.0, 0, 0], [0.1, 0, 0], [0.2, 0, 0]], np.float32)
    w = np.array([0.02, 0.0, 0.02], np.float32)
    adj, off, st, _ = oracle.diagram(pts, w)
    assert st.tolist() == [0, 3, 0]
    assert rows_of(adj, off)[1].size == 0
    assert rows_of(adj, off)[0].tolist() == [2]


def test_duplicates_are_canonical():
    pts = workloads.uniform(200, 4)
    dup = np.concatenate([pts, pts[:5]], 0)
    adj, off, st, _ = oracle.diagram(dup)
    rows = rows_of(adj, off)
    csr_invariants(adj, off, st)
    for k in range(5):
        # a point and its copy have the same row, and never list each other
        assert np.array_equal(rows[k], rows[200 + k])
        assert 200 + k not in rows[k]
        for v in rows[k]:
            assert k in rows[v] and 200 + k in rows[v]


def test_exact_predicate_matches_fractions_random():
    rng = np.random.default_rng(3)
    pts = rng.random((40, 3)).astype(np.float32)
    w = rng.uniform(0, 1e-3, 40).astype(np.float32)
    for _ in range(200):
        idx = rng.choice(40, 5, replace=False)
        seed, tri, q = int(idx[0]), [int(x) for x in idx[1:4]], int(idx[4])
        assert oracle.exact_side(pts, w, seed, tri, q) == _exact_side(pts, w, seed, tri, q)


def test_exact_predicate_near_degenerate():
    # exact ties (five points exactly co-spherical in FP32, or
    # power-co-spherical with weights): P = 0, which the FP64 filter cannot
    # decide, and one float ulp off them; the expansion fallback must match
    # the rationals
    rng = np.random.default_rng(9)
    for trial in range(60):
        pts, w = exact_tie_cloud(rng, weighted=trial % 2 == 1)
        if trial % 3 == 0:
            pts[4] = np.nextafter(pts[4], np.float32(2.0))
        got = oracle.exact_side(pts, w, 0, [1, 2, 3], 4)
        assert got == _exact_side(pts, w, 0, [1, 2, 3], 4)


def test_exact_predicate_box_planes():
    pts = workloads.uniform(10, 1)
    for tri in ([-1, -3, -5], [-2, -4, 1], [-1, 2, 3]):
        for q in (4, 5, 6):
            assert oracle.exact_side(pts, None, 0, tri, q) == _exact_side(pts, None, 0, tri, q)


def test_planted_mistakes_are_caught():
    pts = workloads.uniform(500, 8)
    r = oracle.cells(pts)
    q, _ = qhull_rows(pts)
    # drop one neighbour of cell 7 and flip a status: the comparison must see both
    bad_row = r.row(7)[1:]
    assert not np.array_equal(bad_row, q[7])
    flipped = r.status.copy()
    flipped[3] = 3
    assert not np.array_equal(flipped, np.zeros_like(flipped))


def test_fast_csr_invariants_catch_planted_mistakes():
    """The vectorised CSR check (used on the 20M-point batch) accepts the
    oracle's diagram and rejects planted mistakes: a dropped reverse edge, an
    unsorted row, a self loop, a row for a hidden cell."""
    pts = workloads.uniform(400, 5)
    adj, off, st, _ = oracle.diagram(pts)
    csr_invariants_fast(adj, off, st)
    bad = adj.copy()
    bad[off[10]] = bad[off[10] + 1]  # duplicate in row 10: not strictly increasing, edge missing
    with pytest.raises(AssertionError):
        csr_invariants_fast(bad, off, st)
    bad = adj.copy()
    bad[off[10]], bad[off[10] + 1] = bad[off[10] + 1], bad[off[10]]  # row 10 unsorted
    with pytest.raises(AssertionError):
        csr_invariants_fast(bad, off, st)
    bad = adj.copy()
    bad[off[12]] = 12 if 12 < bad[off[12] + 1] else bad[off[12]]  # self loop in row 12
    with pytest.raises(AssertionError):
        csr_invariants_fast(bad, off, st)
    flipped = st.copy()
    flipped[7] = 3  # hidden, but with a row and listed by its neighbours
    with pytest.raises(AssertionError):
        csr_invariants_fast(adj, off, flipped)
    drop = np.delete(adj, off[20])  # one directed edge gone: its reverse has no partner
    off2 = off.copy()
    off2[21:] -= 1
    with pytest.raises(AssertionError):
        csr_invariants_fast(drop, off2, st)
