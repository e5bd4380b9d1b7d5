"""Shared helpers for the tests (CPU side only)."""
from __future__ import annotations

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
