"""The divergence tool (tests/refsim: the reference's FP32 cell construction
restated on the CPU) classifies the reference's known failure modes, and
agrees with the exact diagram where the reference is exact."""
import numpy as np

from paragram_amd import workloads
from tests import refsim


def test_uniform_cloud_agrees_with_exact_diagram():
    d = refsim.divergence(workloads.uniform(10_000, 0))
    assert d["checked"] == 10_000 and d["differ"] == 0, d


def test_cell_beyond_64_planes_is_a_capacity_status():
    """A point at the centre of 100 points on a sphere has 100 faces: the
    reference's cell (64 planes, 64 vertices) runs out of vertex slots first
    (status 2, voronoi_convex_cell.cuh:21-22, laguerre_ultra.cu:155-171); the
    cause is counted as capacity whatever status the broken cell ends with."""
    rng = np.random.default_rng(1)
    v = rng.normal(size=(100, 3))
    sph = 0.5 + 0.25 * v / np.linalg.norm(v, axis=1, keepdims=True)
    pts = np.concatenate([[[0.5, 0.5, 0.5]], sph]).astype(np.float32)
    d = refsim.divergence(pts, cells=np.array([0]))
    r = d["reasons"]
    assert d["differ"] == 1 and r["status_1_plane_capacity"] + r["status_2_vertex_capacity"] == 1, d


def test_unsigned_too_far_test_misses_a_hidden_cell():
    """A heavy point far from the seed whose power bisector lies behind the
    seed by more than the cell's radius hides the seed (exact status 3); the
    reference's |t| > R test skips it as too far (faster_convex_cell.cuh:391-408)
    and keeps the cell."""
    rng = np.random.default_rng(2)
    near = (0.5 + 0.01 * rng.normal(size=(40, 3))).astype(np.float32)
    pts = np.concatenate([[[0.5, 0.5, 0.5]], near, [[0.9, 0.9, 0.9]]]).astype(np.float32)
    w = np.zeros(pts.shape[0], np.float32)
    w[-1] = 0.6  # d^2 = 0.48: t = (0.48 - 0.6) / (2 d) < 0, and |t| exceeds the small cell's radius
    d = refsim.divergence(pts, w, cells=np.array([0]))
    assert d["differ"] == 1 and d["reasons"]["hidden_mismatch_too_far_behind"] == 1, d
