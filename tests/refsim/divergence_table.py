"""Prints, per configuration, how many cells the reference's own FP32 cell
construction (restated in refsim.cpp) reports differently from the exact
diagram, and why: C1 every cell; C3 and C5 a seeded 20k-cell sample (run
from the repository root: ``python tests/refsim/divergence_table.py``)."""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from paragram_amd import workloads  # noqa: E402
from tests import refsim  # noqa: E402


def main(sample: int = 20_000):
    rows = {}
    t = time.time()
    rows["C1 voronoi 10k seed 0 (every cell)"] = refsim.divergence(workloads.uniform(10_000, 0))
    pts, w = workloads.clustered_power(2_000_000, 2)
    rows[f"C3 power 2M seed 2 ({sample} cells)"] = refsim.divergence(pts, w, workloads.oracle_sample(pts.shape[0], 2, sample))
    pts, w = workloads.learned_power(10_000_000, 4)
    rows[f"C5 power 10M seed 4 ({sample} cells)"] = refsim.divergence(pts, w, workloads.oracle_sample(pts.shape[0], 4, sample))
    for k, v in rows.items():
        print(k, json.dumps(v), flush=True)
    print(f"({time.time() - t:.1f} s)")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 20_000)
