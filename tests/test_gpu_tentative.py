"""Removal tests that tier 0 takes by their FP32 sign, on one cloud per call.

Tier 0 has no FP64 path: a removal test its FP32 filter leaves open is taken
by the sign of its FP32 value and recorded, an FP64 pass after tier 0 decides
every recorded test, and the cell of a wrong sign is rebuilt by the
FP64-capable instance (or, when only a few cells are flagged, by tier m).
A LiDAR-like scan is full of such tests (near-coplanar ground points), so one
cloud of it, called alone (not batched), must take tentative tests, find some
signs wrong, and still match the oracle in every cell.
"""
import numpy as np
import pytest

import oracle
from paragram_amd import workloads
from tests.helpers import csr_invariants

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_lidar_cloud_tentative_tests_are_exact(gpu):
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    clouds = workloads.lidar_batch(3, n_clouds=16)
    cloud = min(clouds, key=lambda c: c.shape[0])
    log(f"one LiDAR-like cloud of {cloud.shape[0]} points (seed 3, smallest of 16), one call")
    ins = _engine.Instrument()
    a, o, s = _engine.build_diagram(_C.lib, torch.from_numpy(cloud).to(gpu), None, instrument=ins)
    torch.cuda.synchronize()
    st = ins.stats
    log(f"stats: tentative {st['tentative']}, wrong {st['tentative_wrong']}, "
        f"to the FP64 tier directly {st['bail_test']}, rebuilt cells logged {st['rebuilt_cells']}")
    assert st["tentative"] > 0, "no removal test was taken by its FP32 sign"
    assert st["unresolved"] == 0 and st["asymmetric"] == 0 and st["dropped"] == 0
    assert st["tentative_wrong"] <= st["tentative"]
    got = (a.cpu().numpy(), o.cpu().numpy(), s.cpu().numpy())
    csr_invariants(*got)
    adj, off, status, _ = oracle.diagram(cloud, None)
    assert np.array_equal(got[2], status), f"status differs in {int((got[2] != status).sum())} cells"
    assert np.array_equal(got[1], off), f"offsets differ ({int((np.diff(got[1]) != np.diff(off)).sum())} rows)"
    assert np.array_equal(got[0], adj), "adjacency differs"
