"""The batched LiDAR configuration at its real size: 256 clouds (seed 3,
log-uniform 20k-200k points, about 20M in all) in one launch.

Checked against the oracle: a seeded sample of 400 cells of every cloud and
every cell of the four largest clouds (cloud-local ids, rows and statuses
bit-exact); CSR invariants of every cloud; no cell truncated (status 1) and
no row-pool shortfall; every cell whose removal tests needed expansion
arithmetic checked against the oracle as well.  Tier 0 takes the removal
tests its FP32 filter leaves open by their FP32 sign and the FP64 check
after it sends the cells of wrong signs to the FP64 twin: both happen here,
and the first cells rebuilt that way are checked against the oracle too.
"""
import concurrent.futures

import numpy as np
import pytest

import oracle
from paragram_amd import workloads
from tests.helpers import csr_invariants_fast

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_c4_256_lidar_clouds_one_launch(gpu):
    import torch

    from paragram_amd import _engine
    from paragram_amd._backend import _C

    log("C4: generating 256 LiDAR clouds (seed 3)")
    clouds = workloads.lidar_batch(3)
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in clouds])]).astype(np.int64)
    assert len(clouds) == 256
    pts = torch.from_numpy(np.concatenate(clouds)).to(gpu)
    log(f"C4: one batched launch over {int(off[-1]):,} points")
    ins = _engine.Instrument()
    a, o, s = _engine.build_diagram(_C.lib, pts, None, clouds=off, instrument=ins)
    parts = [tuple(t.cpu().numpy() for t in r) for r in _engine.split_batch(a, o, s, off)]
    del a, o, s, pts
    st = ins.stats
    log(f"C4: stats {st}")
    assert st["truncated"] == 0, "a cell was cut short (status 1)"
    assert st["pool_short"] == 0
    assert st["unresolved"] == 0, "a removal test was left to the FP64 sign"
    assert st["asymmetric"] == 0 and st["dropped"] == 0
    assert st["tentative"] > 0 and 0 < st["tentative_wrong"] < st["tentative"]
    assert st["rebuilt_cells"], "the first rebuilt cells are logged"

    log("C4: CSR invariants of every cloud")
    with concurrent.futures.ThreadPoolExecutor(16) as ex:
        list(ex.map(lambda p: csr_invariants_fast(*p), parts))

    sizes = off[1:] - off[:-1]
    full = set(np.argsort(sizes)[-4:].tolist())
    exact_cells = st["exact_cells"] + st["rebuilt_cells"]
    log(f"C4: oracle on 400 cells per cloud and every cell of clouds {sorted(full)}; "
        f"cells with expansion-arithmetic tests, then the first cells rebuilt after a wrong FP32 sign: "
        f"{exact_cells}")
    bad = []
    checked = 0
    for c, cloud in enumerate(clouds):
        adj, offs, status = parts[c]
        n_c = cloud.shape[0]
        if c in full:
            sample = np.arange(n_c, dtype=np.int64)
        else:
            sample = workloads.oracle_sample(n_c, 3 + c, 400)
        extra = [g - off[c] for g in exact_cells if off[c] <= g < off[c + 1]]
        if extra:
            sample = np.unique(np.concatenate([sample, np.asarray(extra, np.int64)]))
        ref = oracle.cells(cloud, None, sample)
        for k, i in enumerate(sample):
            if status[i] != ref.status[k] or not np.array_equal(adj[offs[i]:offs[i + 1]], ref.row(k)):
                bad.append((c, int(i)))
        checked += sample.shape[0]
    log(f"C4: {checked} cells checked, {len(bad)} differ")
    assert not bad, f"{len(bad)} of {checked} cells differ from the oracle, first {bad[:5]}"
