"""Benchmark of the hot path: paragram_amd.voronoi_diagram on one MI355X.

Workload (the configuration the headline metric is quoted on): plain Voronoi
of 1,000,000 float32 points i.i.d. uniform in [0,1)^3 (seed 1).  One step is
one full call, points already resident in HBM to Diagram out: index build,
cell construction, CSR assembly, including its one device-to-host read.

Prints one JSON line.  ``value`` is Mcells/s = N * steps / timed seconds / 1e6.
``roofline`` prices the dominant kernel (cell construction) at the
algorithmic bytes per cell B = 24 + 20 F (F = E / N, measured): the seed's
float4, each true neighbour's float4 once, the neighbour id written, plus
offset and status; achieved = N B / (mean cell-kernel duration from HIP
events on the launch stream), peak 8 TB/s.  ``cpu_baseline`` times the CPU
oracle (FP64 filter + exact fallback, OpenMP) on a seeded sample of cells of
the same cloud, and the same sample gives ``exact_cells``: the fraction of
sampled GPU rows and statuses equal to the oracle's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcells/s @1M & 10M pts, plain vs learned-radius form; % cells exact vs FP64"


def _traffic_record(workload_key: str):
    """HBM bytes per cell-kernel launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "cell_kernel_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec.get(workload_key)
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=20_000, help="cells timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU sharding is "
              "described in DESIGN.md and not built)" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import paragram_amd
    from paragram_amd import _backend, _engine, workloads

    if _backend._C is None:
        raise RuntimeError("HIP library not built: " + str(_backend._load_error))
    dev = torch.device("cuda:0")
    pts_np = workloads.uniform(args.n, args.seed)
    pts = torch.from_numpy(pts_np).to(dev)
    lib = _backend._C.lib

    def step(instr=None):
        return _engine.build_diagram(lib, pts, None, 8, False, None, None, -1, instrument=instr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    instr = _engine.Instrument()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step(instr)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = args.n * args.steps / elapsed / 1e6

    adj, off, status = (t.cpu().numpy() for t in out)
    e_total = int(off[-1])
    f_bar = e_total / args.n
    b_cell = 24.0 + 20.0 * f_bar
    kern_ms = [a.elapsed_time(b) for a, b in instr.cell_events]
    kern_s = float(np.mean(kern_ms)) / 1e3
    achieved = args.n * b_cell / kern_s / 1e9
    wkey = f"uniform_{args.n}_seed{args.seed}"
    traffic = _traffic_record(wkey)

    cpu = None
    exact = None
    if not args.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        sample = workloads.oracle_sample(args.n, args.seed, args.cpu_sample)
        # the oracle builds its k-d tree per call: time that separately
        tb = time.perf_counter()
        oracle.cells(pts_np, None, sample[:1], nthreads=cores)
        t_build = time.perf_counter() - tb
        tc = time.perf_counter()
        ref = oracle.cells(pts_np, None, sample, nthreads=cores)
        t_cells = time.perf_counter() - tc - t_build
        cpu = {"value": round(sample.shape[0] / max(t_cells, 1e-9) / 1e6, 4), "unit": "Mcells/s",
               "cores": cores, "kind": "port",
               "sample": f"{sample.shape[0]} seeded cells of the same {args.n}-point cloud; CPU oracle "
                         "(FP64 filtered predicates, Shewchuk-expansion exact fallback), OpenMP; k-d tree build "
                         f"({t_build:.2f}s) excluded"}
        ok = 0
        for k, i in enumerate(sample):
            g = adj[off[i]:off[i + 1]]
            ok += int(status[i] == ref.status[k] and np.array_equal(g, ref.row(k)))
        exact = {"checked": int(sample.shape[0]), "exact": ok, "frac": ok / sample.shape[0]}

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain Voronoi, {args.n:,} points fp32 i.i.d. U[0,1)^3, seed {args.seed}",
                   "form": "plain", "n_points": args.n, "edges": e_total, "mean_faces": round(f_bar, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic,
                     "kernel": "pg::cell_kernel", "kernel_ms": round(kern_s * 1e3, 4),
                     "bytes_per_cell": round(b_cell, 2)},
        "cpu_baseline": cpu,
        "exact_cells": exact,
        "stats": instr.stats,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
