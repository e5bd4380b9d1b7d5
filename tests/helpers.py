"""Shared helpers for the tests (CPU side only)."""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def rows_of(adjacency, offsets):
    adjacency = np.asarray(adjacency)
    offsets = np.asarray(offsets)
    return [adjacency[offsets[i]:offsets[i + 1]] for i in range(offsets.shape[0] - 1)]


def csr_invariants(adjacency, offsets, status):
    """CSR invariants of the reference's output (README.md:65-69, adjacency.cuh:78-159)."""
    adjacency = np.asarray(adjacency)
    offsets = np.asarray(offsets)
    status = np.asarray(status)
    n = status.shape[0]
    assert offsets.shape == (n + 1,)
    assert offsets[0] == 0 and offsets[-1] == adjacency.shape[0]
    assert np.all(np.diff(offsets) >= 0)
    rows = rows_of(adjacency, offsets)
    edges = set()
    for i, r in enumerate(rows):
        assert np.all(np.diff(r) > 0), f"row {i} not strictly increasing"
        assert not np.any(r == i), f"row {i} has a self loop"
        assert np.all((r >= 0) & (r < n))
        if status[i] == 3:
            assert r.size == 0
        for j in r:
            edges.add((i, int(j)))
    for (i, j) in edges:
        assert (j, i) in edges, f"edge {i}->{j} has no reverse"
        assert status[j] != 3


def exact_side_fractions(points, weights, seed, tri, q, box=None):
    """(sign(P) sign(D), sign(D), sign(P)) of the removal predicate in rational
    arithmetic: the vertex of planes tri of the seed's cell against plane q
    (ids >= 0 points, -1-(2*axis+side) box planes of the FP32 box)."""
    pts = np.asarray(points, np.float32)
    w = np.zeros(len(pts), np.float32) if weights is None else np.asarray(weights, np.float32)
    if box is None:
        import oracle

        box = oracle.box(pts)
    lo, hi = box
    s = [Fraction(float(v)) for v in pts[seed]]
    ws = Fraction(float(w[seed]))

    def plane(o):
        if o >= 0:
            a = [Fraction(float(pts[o][k])) - s[k] for k in range(3)]
            return a, sum(x * x for x in a) - (Fraction(float(w[o])) - ws)
        code = -1 - o
        axis, side = code // 2, code % 2
        a = [Fraction(0)] * 3
        if side == 0:
            a[axis] = Fraction(-1)
            return a, 2 * (s[axis] - Fraction(float(lo[axis])))
        a[axis] = Fraction(1)
        return a, 2 * (Fraction(float(hi[axis])) - s[axis])

    (a1, h1), (a2, h2), (a3, h3) = (plane(t) for t in tri)
    aq, hq = plane(q)

    def cross(u, v):
        return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]

    c1, c2, c3 = cross(a2, a3), cross(a3, a1), cross(a1, a2)
    D = sum(x * y for x, y in zip(a1, c1))
    N = [h1 * c1[k] + h2 * c2[k] + h3 * c3[k] for k in range(3)]
    P = sum(x * y for x, y in zip(aq, N)) - hq * D

    def sgn(x):
        return (x > 0) - (x < 0)

    return sgn(P) * sgn(D), sgn(D), sgn(P)


def sphere_lattice_points(r2: int):
    """Integer points (x, y, z) with x^2 + y^2 + z^2 = r2."""
    r = int(np.floor(np.sqrt(r2))) + 1
    out = []
    for x in range(-r, r + 1):
        for y in range(-r, r + 1):
            z2 = r2 - x * x - y * y
            if z2 < 0:
                continue
            z = int(round(np.sqrt(z2)))
            if z * z == z2:
                out.append((x, y, z))
                if z:
                    out.append((x, y, -z))
    return np.array(out, np.int64)


def exact_tie_cloud(rng, k=5, weighted=False):
    """k points that are exactly co-spherical in FP32 (power-co-spherical when
    weighted): every removal test among them is an exact tie, P = 0, which no
    floating-point filter can decide.  Integer points on x^2+y^2+z^2 = 81
    (or on several spheres, with weights w = |p|^2 - 81 making the power
    distance to the centre equal), scaled by 1/32 about (0.5, 0.5, 0.5): all
    exact in float32."""
    if weighted:
        cand = np.concatenate([sphere_lattice_points(r2) for r2 in (81, 89, 98, 101)])
    else:
        cand = sphere_lattice_points(81)
    idx = rng.choice(cand.shape[0], k, replace=False)
    ip = cand[idx]
    pts = (0.5 + ip / 32.0).astype(np.float32)
    assert np.array_equal(((pts.astype(np.float64) - 0.5) * 32.0), ip)
    if not weighted:
        return pts, None
    w = ((ip * ip).sum(1) - 81) / 1024.0  # |x - c|^2 - w = 81 / 1024 for every point
    return pts, w.astype(np.float32)


def csr_invariants_fast(adjacency, offsets, status):
    """csr_invariants, vectorised for large diagrams: offsets well formed, ids
    in range, rows strictly increasing without self loops, hidden cells with
    empty rows and never listed, every edge with its reverse."""
    adj = np.asarray(adjacency).astype(np.int64)
    off = np.asarray(offsets).astype(np.int64)
    st = np.asarray(status)
    n = st.shape[0]
    assert off.shape == (n + 1,)
    assert off[0] == 0 and off[-1] == adj.shape[0]
    deg = np.diff(off)
    assert np.all(deg >= 0)
    assert np.all((adj >= 0) & (adj < n))
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    assert np.all(adj != rows), "self loop"
    keys = rows * n + adj
    assert np.all(np.diff(keys) > 0), "a row is not strictly increasing"
    assert np.all(deg[st == 3] == 0), "a hidden cell has a row"
    assert np.all(st[adj] != 3), "a hidden cell is listed as a neighbour"
    assert np.array_equal(np.sort(adj * n + rows), keys), "an edge has no reverse"
