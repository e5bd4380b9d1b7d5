"""Writes the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no fixtures (no tests at all), so these are generated
here and pinned independently before they are written:

* diagrams: the CPU oracle's rows, each checked against Qhull (Delaunay of the
  points plus their mirrors across the six box planes; for weights, the lower
  hull of the lifted points) -- the same construction tests/test_oracle.py
  uses -- and written only if every row and status agrees;
* the cubic lattice: face adjacency in closed form (every removal test of a
  diagonal neighbour is an exact tie, see tests/test_gpu_parity.py);
* predicates: signs of the removal test in rational arithmetic (Fractions),
  on exact ties (integer points on a common sphere), their one-ulp
  neighbours and random configurations.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from paragram_amd import workloads  # noqa: E402
from tests.helpers import exact_side_fractions, exact_tie_cloud  # noqa: E402
from tests.test_oracle import qhull_rows  # noqa: E402


def diagram_fixture(name, pts, w=None):
    adj, off, st, _ = oracle.diagram(pts, w)
    q, present = qhull_rows(pts, w)
    for i in range(pts.shape[0]):
        assert np.array_equal(adj[off[i]:off[i + 1]], q[i]), f"{name}: row {i} differs from Qhull"
        assert (st[i] == 3) == (not present[i]), f"{name}: status {i} differs from Qhull"
    arrays = {"points": pts, "adjacency": adj, "offsets": off, "status": st}
    if w is not None:
        arrays["weights"] = w
    np.savez(os.path.join(HERE, name + ".npz"), **arrays)
    print(f"{name}: {pts.shape[0]} points, {adj.shape[0]} entries, {(st == 3).sum()} hidden")


def lattice_fixture(m=4):
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) / 16.0
    rows = []
    for x in range(m):
        for y in range(m):
            for z in range(m):
                nb = []
                for d in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    u, v, t = x + d[0], y + d[1], z + d[2]
                    if 0 <= u < m and 0 <= v < m and 0 <= t < m:
                        nb.append((u * m + v) * m + t)
                rows.append(sorted(nb))
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    adj = np.concatenate(rows).astype(np.int32)
    np.savez(os.path.join(HERE, f"lattice_{m}.npz"), points=g, adjacency=adj, offsets=off,
             status=np.zeros(g.shape[0], np.int32))
    print(f"lattice_{m}: face adjacency of {g.shape[0]} lattice points")


def predicate_fixture(count=120):
    rng = np.random.default_rng(2024)
    P, W, E = [], [], []
    for t in range(count):
        kind = t % 4
        if kind == 3:
            pts = rng.random((5, 3)).astype(np.float32)
            w = rng.uniform(0, 1e-2, 5).astype(np.float32)
        else:
            pts, w = exact_tie_cloud(rng, weighted=kind == 1)
            if w is None:
                w = np.zeros(5, np.float32)
            if kind == 2:
                pts[4] = np.nextafter(pts[4], np.float32(2.0))
        P.append(pts)
        W.append(w)
        E.append(exact_side_fractions(pts, w, 0, [1, 2, 3], 4))
    E = np.array(E, np.int32)
    np.savez(os.path.join(HERE, "predicates.npz"), points=np.array(P), weights=np.array(W), expected=E)
    print(f"predicates: {count} queries (seed 0, vertex planes 1 2 3, query plane 4), "
          f"{int((E[:, 2] == 0).sum())} exact ties")


def main():
    diagram_fixture("voronoi_uniform_300", workloads.uniform(300, 5))
    rng = np.random.default_rng(6)
    pts = rng.random((400, 3)).astype(np.float32)
    diagram_fixture("power_hidden_400", pts, rng.uniform(0, 1.2e-2, 400).astype(np.float32))
    cp, cw = workloads.clustered_power(500, 8)
    diagram_fixture("power_clustered_500", cp, cw)
    lattice_fixture(4)
    predicate_fixture()


if __name__ == "__main__":
    main()
