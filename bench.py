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
events on the launch stream), peak 8 TB/s.  ``cpu_baseline`` times the same
algorithm on the host cores (paragram_amd.cpu, C++/OpenMP, all cores of the
process's affinity set) over the whole cloud; ``exact_cells`` compares a
seeded sample of GPU rows and statuses with the CPU oracle (FP64 filter +
exact fallback, k-d tree).

``extra`` (skipped with --no-extra) measures the rest of the metric in the
same run: the 10M-point power configuration in the plain form and in the
learned-radius form (the committed network), their ratio, the 256-cloud
LiDAR batch in one launch, and the 10k and 2M configurations; each entry has
its ms per call (median of the timed calls), Mcells/s, FP64-recheck and
exact-arithmetic counts, and an oracle sample (every cell up to 1M points, a
seeded 100k-cell sample at 10M, 200 cells of each of 32 clouds of the batch).
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


# name -> (generator in paragram_amd.workloads, default n, default seed, description)
WORKLOADS = {
    "c1": ("uniform", 10_000, 0, "plain Voronoi, {n:,} points fp32 i.i.d. U[0,1)^3, seed {seed}"),
    "c2": ("uniform", 1_000_000, 1, "plain Voronoi, {n:,} points fp32 i.i.d. U[0,1)^3, seed {seed}"),
    "c3": ("clustered_power", 2_000_000, 2, "power diagram, {n:,} points fp32: 90% from 64 Gaussians "
           "(sigma 0.02, clamped) + 10% uniform, weights U[0, 3e-5], seed {seed}"),
    "c5": ("learned_power", 10_000_000, 4, "power diagram, {n:,} points fp32: 50% uniform + 50% from 256 "
           "Gaussians (sigma 0.01), weights U[0, 5e-6], seed {seed} (plain form)"),
    "c4": ("lidar_batch", 256, 3, "batched Voronoi, {n} LiDAR-like clouds (log-uniform 20k-200k points, "
           "{total:,} points in all), seed {seed}, one launch"),
}
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "cell_kernel_traffic.json")
CELL_KERNELS = ("cell_group_kernel", "cell_tentative_check_kernel", "cell_group_exact_kernel",
                "cell_group_exact_batch_kernel", "cell_group_mid_kernel", "cell_group_big_kernel",
                "cell_group_huge_kernel")  # (the FP64 twin's batch form launches for batches)
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _traffic_record(workload_key: str):
    """HBM bytes per cell-construction launch from the committed rocprofv3 PMC summary, if any."""
    try:
        with open(TRAFFIC_PATH) as f:
            rec = json.load(f)
        return rec.get(workload_key)
    except (OSError, ValueError):
        return None


def _pmc_values(run_dir: str, counter: str):
    """{dispatch id: (kernel name, value)} of one counter from a rocprofv3 --pmc csv run."""
    import csv
    import glob

    out = {}
    for path in glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    out[int(r["Dispatch_Id"])] = (r["Kernel_Name"], float(r["Counter_Value"]))
    return out


def summarize_pmc(fetch_dir: str, write_dir: str, copy_mib: int, workload_key: str, command: str) -> dict:
    """Cell-construction HBM bytes per launch from two --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per
    dispatch), each scaled by the same counter's reading for a device copy of copy_mib MiB."""
    res = {}
    for counter, run_dir in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        vals = _pmc_values(run_dir, counter)
        copies = [v for name, v in vals.values() if name.startswith(COPY_KERNEL)]
        if not copies:
            raise RuntimeError(f"no {COPY_KERNEL} dispatch in {run_dir}")
        factor = max(copies) / (copy_mib * 1024.0)  # counter KiB per byte-KiB moved by the copy
        per_kernel = {}
        for name, v in vals.values():
            short = name.split("(")[0].split("::")[-1]
            if short in CELL_KERNELS:
                per_kernel.setdefault(short, []).append(v)
        launches = len(per_kernel.get(CELL_KERNELS[0], []))
        kib = sum(sum(v) for v in per_kernel.values()) / max(launches, 1)
        res[counter] = {"kib_per_launch_raw": round(kib, 1), "copy_factor": round(factor, 4),
                        "bytes_per_launch": round(kib * 1024.0 / factor)}
    total = res["FETCH_SIZE"]["bytes_per_launch"] + res["WRITE_SIZE"]["bytes_per_launch"]
    rec = {"bytes_per_launch": total, "fetch": res["FETCH_SIZE"], "write": res["WRITE_SIZE"],
           "kernels": list(CELL_KERNELS), "copy_mib": copy_mib, "command": command}
    try:
        with open(TRAFFIC_PATH) as f:
            allrec = json.load(f)
    except (OSError, ValueError):
        allrec = {}
    allrec[workload_key] = rec
    with open(TRAFFIC_PATH, "w") as f:
        json.dump(allrec, f, indent=1)
        f.write("\n")
    return rec


STAT_KEYS = ("fp64_rechecks", "exact", "unresolved", "promoted", "truncated", "pool_short", "edges", "tentative",
             "tentative_wrong", "bail_test")


def _progress(msg: str) -> None:
    """One line per finished phase on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _time_calls(fn, warmup: int, steps: int):
    """Median and all timings (ms) of steps calls after warmup calls, each
    bracketed by device synchronisations."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ms = []
    res = None
    for _ in range(steps):
        t = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t) * 1e3)
    return float(sorted(ms)[len(ms) // 2]), [round(x, 3) for x in ms], res


def _sample_check(adj, off, status, pts_np, w_np, sample):
    import numpy as np

    import oracle

    ref = oracle.cells(pts_np, w_np, sample)
    ok = 0
    for k, i in enumerate(sample):
        ok += int(status[i] == ref.status[k] and np.array_equal(adj[off[i]:off[i + 1]], ref.row(k)))
    return ok


def measure_extra(lib, dev, cpu_check: bool = True) -> dict:
    """The other configurations of the metric, timed in this run (see the module docstring)."""
    import numpy as np
    import torch

    from paragram_amd import _engine, learned, workloads

    res = {}

    def entry(desc, n, ms, all_ms, instr, exact):
        st = instr.stats or {}
        e = {"workload": desc, "n_points": n, "ms_per_call": round(ms, 3), "calls_ms": all_ms,
             "mcells_s": round(n / ms / 1e3, 3), "stats": {k: st.get(k) for k in STAT_KEYS}}
        if exact is not None:
            e["exact_cells"] = exact
        return e

    def single(key, gen_args, weights_key=None, steps=3, model=None, sample_count=None):
        gen, n, seed, desc = WORKLOADS[key]
        made = getattr(workloads, gen)(*gen_args)
        pts_np, w_np = made if isinstance(made, tuple) else (made, None)
        pts = torch.from_numpy(pts_np).to(dev)
        w = None if w_np is None else torch.from_numpy(w_np).to(dev).reshape(-1, 1)
        instr = _engine.Instrument()

        def call():
            instr.cell_events.clear()
            return _engine.build_diagram(lib, pts, w, 8, False, None, None, -1, instrument=instr, radius_model=model)

        ms, all_ms, out = _time_calls(call, 1, steps)
        exact = None
        if cpu_check:
            adj, off, status = (t.cpu().numpy() for t in out)
            sample = workloads.oracle_sample(pts_np.shape[0], seed, sample_count or pts_np.shape[0])
            ok = _sample_check(adj, off, status, pts_np, w_np, sample)
            exact = {"checked": int(sample.shape[0]), "exact": ok, "frac": ok / sample.shape[0]}
        del out, pts, w
        torch.cuda.empty_cache()
        _progress(f"extra {key}{' learned' if model is not None else ''}: {ms:.2f} ms per call")
        return desc.format(n=n, seed=seed, total=pts_np.shape[0]), pts_np.shape[0], ms, all_ms, instr, exact

    res["c1"] = entry(*single("c1", (10_000, 0), steps=5))
    res["c3"] = entry(*single("c3", (2_000_000, 2), steps=3, sample_count=20_000))
    model = learned.load()
    plain = entry(*single("c5", (10_000_000, 4), steps=3, sample_count=100_000))
    desc, n, ms, all_ms, instr, exact = single("c5", (10_000_000, 4), steps=3, model=model, sample_count=100_000)
    lrn = entry(desc.replace("(plain form)", "(learned-radius form: committed radius_mlp.npz)"), n, ms, all_ms, instr,
                exact)
    res["c5"] = {"plain": plain, "learned": lrn,
                 "learned_over_plain": round(plain["ms_per_call"] / lrn["ms_per_call"], 4)}

    # the LiDAR batch: one launch for all 256 clouds
    gen, n_clouds, seed, desc = WORKLOADS["c4"]
    cl = workloads.lidar_batch(seed, n_clouds=n_clouds)
    clouds = np.concatenate([[0], np.cumsum([c.shape[0] for c in cl])]).astype(np.int64)
    pts = torch.from_numpy(np.concatenate(cl)).to(dev)
    instr = _engine.Instrument()

    def batch_call():
        instr.cell_events.clear()
        out = _engine.build_diagram(lib, pts, None, 8, False, None, None, -1, instrument=instr, clouds=clouds)
        return _engine.split_batch(*out, clouds)

    ms, all_ms, parts = _time_calls(batch_call, 1, 3)
    # the learned + batched form: the same batch started from the learned radius
    instr_l = _engine.Instrument()

    def batch_learned():
        instr_l.cell_events.clear()
        out = _engine.build_diagram(lib, pts, None, 8, False, None, None, -1, instrument=instr_l, clouds=clouds,
                                    radius_model=model)
        return _engine.split_batch(*out, clouds)

    ms_l, all_ms_l, _ = _time_calls(batch_learned, 1, 3)
    _progress(f"extra c4 learned: {ms_l:.2f} ms per batch")
    exact = None
    if cpu_check:
        ok = checked = 0
        for c in range(0, n_clouds, n_clouds // 32):
            a, o, s = (t.cpu().numpy() for t in parts[c])
            sample = workloads.oracle_sample(cl[c].shape[0], seed + c, 200)
            ok += _sample_check(a, o, s, cl[c], None, sample)
            checked += sample.shape[0]
        exact = {"checked": checked, "exact": ok, "frac": ok / checked}
    res["c4"] = entry(desc.format(n=n_clouds, seed=seed, total=int(clouds[-1])), int(clouds[-1]), ms, all_ms, instr,
                      exact)
    res["c4_learned"] = entry(desc.format(n=n_clouds, seed=seed, total=int(clouds[-1]))
                              + " (learned-radius form: committed radius_mlp.npz)", int(clouds[-1]), ms_l, all_ms_l,
                              instr_l, None)
    res["c4_learned"]["learned_over_plain"] = round(ms / ms_l, 4)
    _progress(f"extra c4: {ms:.2f} ms per batch")
    return res


def radius_bound(n=None) -> dict:
    """What a starting radius can gain on the 10M power configuration: the same diagram built
    from the kNN pre-clip (plain), from a perfect cube (every cell's exact final half-width, read
    back from a first build) with and without the pre-clip, from the committed network, and from
    the box alone; ms per call (median of 3 after one warm-up) and cell-kernel ms.  Every variant
    is checked byte-identical to the plain output."""
    import torch

    from paragram_amd import _backend, _engine, learned, workloads

    lib = _backend._C.lib
    gen, n_def, seed, desc = WORKLOADS["c5"]
    n = n_def if n is None else n
    pts_np, w_np = workloads.learned_power(n, seed)
    dev = torch.device("cuda:0")
    P = torch.from_numpy(pts_np).to(dev)
    W = torch.from_numpy(w_np).to(dev).reshape(-1, 1)
    ext = torch.zeros(n, dtype=torch.float32, device=dev)
    ref = _engine.build_diagram(lib, P, W, extent=ext)
    perfect = (ext * 1.001 + 1e-7).contiguous()
    out = {"workload": desc.format(n=n, seed=seed, total=n)}
    for label, kw in (("plain (kNN-16 pre-clip)", {}), ("perfect cube + kNN pre-clip", {"init_h": perfect}),
                      ("learned network + kNN pre-clip", {"radius_model": learned.load()}),
                      ("box start, no pre-clip", {"knn": 0}), ("perfect cube, no pre-clip", {"knn": 0, "init_h": perfect})):
        ins = _engine.Instrument()
        knn = kw.pop("knn", 8)

        def call():
            ins.cell_events.clear()
            return _engine.build_diagram(lib, P, W, knn, instrument=ins, **kw)

        ms, all_ms, res = _time_calls(call, 1, 3)
        same = all(bool(torch.equal(x, y)) for x, y in zip(res, ref))
        if not same:
            raise RuntimeError(f"radius bound: {label} differs from the plain output")
        cell_ms = sum(a.elapsed_time(b) for a, b in ins.cell_events)
        out[label] = {"ms_per_call": round(ms, 2), "cell_ms": round(cell_ms, 2), "clips": ins.stats["clips"],
                      "candidates": ins.stats["candidates"], "promoted": ins.stats["promoted"]}
        _progress(f"radius bound, {label}: {ms:.2f} ms")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2",
                    help="c2 (default, the headline configuration), or another configuration of the metric")
    ap.add_argument("--n", type=int, default=None, help="override the workload's point count")
    ap.add_argument("--seed", type=int, default=None, help="override the workload's seed")
    ap.add_argument("--cpu-sample", type=int, default=20_000, help="cells timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--form", choices=("plain", "learned"), default="plain",
                    help="learned: start each cell from the committed radius network's prediction")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed steps, one device copy of this many MiB (PMC calibration)")
    ap.add_argument("--radius-bound", action="store_true",
                    help="time the 10M power configuration from a perfect starting radius (each cell's exact "
                         "final extent) against the plain, learned and box starts, print one JSON line and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this command "
                         "with --copy-mib into profiles/cell_kernel_traffic.json and exit")
    args = ap.parse_args()
    gen, n_def, seed_def, desc = WORKLOADS[args.workload]
    args.n = n_def if args.n is None else args.n
    args.seed = seed_def if args.seed is None else args.seed
    wkey = f"{gen}_{args.n}_seed{args.seed}"
    if args.pmc_summary:
        if args.copy_mib <= 0:
            print("bench.py: --pmc-summary needs the --copy-mib of the profiled runs", file=sys.stderr)
            return 2
        cmd = ("rocprofv3 --pmc {FETCH_SIZE|WRITE_SIZE} --output-format csv -d <dir> -o run -- "
               f"python bench.py --n {args.n} --seed {args.seed} --steps {args.steps} --warmup {args.warmup} "
               f"--no-cpu --copy-mib {args.copy_mib}")
        rec = summarize_pmc(args.pmc_summary[0], args.pmc_summary[1], args.copy_mib, wkey, cmd)
        print(json.dumps(rec))
        return 0
    if args.radius_bound:
        print(json.dumps({"radius_bound": radius_bound(args.n if args.workload == "c5" else None)}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (multi-GPU sharding is "
              "designed but not built)" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import paragram_amd
    from paragram_amd import _backend, _engine, workloads

    if _backend._C is None:
        raise RuntimeError("HIP library not built: " + str(_backend._load_error))
    dev = torch.device("cuda:0")
    clouds = None
    if gen == "lidar_batch":
        cl = workloads.lidar_batch(args.seed, n_clouds=args.n)
        clouds = np.concatenate([[0], np.cumsum([c.shape[0] for c in cl])]).astype(np.int64)
        made = np.concatenate(cl)
        del cl
    else:
        made = getattr(workloads, gen)(args.n, args.seed)
    pts_np, w_np = made if isinstance(made, tuple) else (made, None)
    n_cells = pts_np.shape[0]
    pts = torch.from_numpy(pts_np).to(dev)
    w = None if w_np is None else torch.from_numpy(w_np).to(dev).reshape(-1, 1)
    lib = _backend._C.lib

    model = None
    if args.form == "learned":
        from paragram_amd import learned

        model = learned.load()

    def step(instr=None):
        out = _engine.build_diagram(lib, pts, w, 8, False, None, None, -1, instrument=instr, clouds=clouds,
                                    radius_model=model)
        if clouds is not None:  # the API's per-cloud results (cloud-local ids)
            _engine.split_batch(*out, clouds)
        return out

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
    value = n_cells * args.steps / elapsed / 1e6
    _progress(f"{args.steps} timed steps: {ms_per_step:.3f} ms per step")
    per_cloud_ms = None
    if clouds is not None:  # the same clouds as one call each, for comparison
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for c in range(clouds.shape[0] - 1):
            _engine.build_diagram(lib, pts[clouds[c]:clouds[c + 1]], None, 8, False, None, None, -1)
        torch.cuda.synchronize()
        per_cloud_ms = (time.perf_counter() - tp) * 1e3
        _progress(f"the clouds one call each: {per_cloud_ms:.1f} ms")

    if args.copy_mib > 0:  # a plain copy of known size for the PMC byte counters
        src = torch.ones(args.copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    adj, off, status = (t.cpu().numpy() for t in out)
    e_total = int(off[-1])
    f_bar = e_total / n_cells
    b_cell = 24.0 + 20.0 * f_bar
    kern_ms = [a.elapsed_time(b) for a, b in instr.cell_events]
    kern_s = float(np.mean(kern_ms)) / 1e3
    achieved = n_cells * b_cell / kern_s / 1e9
    trec = _traffic_record(wkey)
    traffic = trec["bytes_per_launch"] if trec else None

    cpu = None
    exact = None
    if not args.no_cpu:
        import oracle
        from paragram_amd import cpu as host

        cores = len(os.sched_getaffinity(0))
        if clouds is not None:  # one cloud of the batch on the host, and the oracle on it
            pts_np, w_np = pts_np[clouds[0]:clouds[1]], None
        # the same clipping algorithm on the host cores, over the whole cloud
        hl = host.CpuLibrary()
        hl.diagram(pts_np[:2000], None if w_np is None else w_np[:2000], nthreads=cores)  # thread start-up
        th = time.perf_counter()
        h_adj, h_off, h_st, h_stats = hl.diagram(pts_np, w_np, nthreads=cores)
        t_host = time.perf_counter() - th
        _progress(f"host-core build: {t_host:.2f} s on {cores} cores")
        same = None
        if clouds is None:
            same = bool(np.array_equal(h_adj, adj) and np.array_equal(h_off, off) and np.array_equal(h_st, status))
        cpu = {"value": round(pts_np.shape[0] / t_host / 1e6, 4), "unit": "Mcells/s", "cores": cores, "kind": "port",
               "sample": f"the whole {pts_np.shape[0]}-point cloud (index, cells and CSR, {t_host:.2f} s) with the "
                         "GPU kernels' algorithm on the host cores (paragram_amd/csrc/pg_cpu.cpp: filtered FP32, FP64 "
                         "recheck, exact fallback, kNN pre-clip, same prunes), C++/OpenMP; "
                         + ("output identical to the GPU's" if same else
                            "output differs from the GPU's" if same is not None else "cloud 0 of the batch")}
        # the oracle (FP64 filter + exact fallback, k-d tree) as the checker on a seeded sample
        sample = workloads.oracle_sample(pts_np.shape[0], args.seed, args.cpu_sample)
        ref = oracle.cells(pts_np, w_np, sample, nthreads=cores)
        ok = 0
        for k, i in enumerate(sample):  # (cloud 0 of a batch: its ids are the global ids)
            g = adj[off[i]:off[i + 1]]
            ok += int(status[i] == ref.status[k] and np.array_equal(g, ref.row(k)))
        exact = {"checked": int(sample.shape[0]), "exact": ok, "frac": ok / sample.shape[0]}

    extra = None
    if not args.no_extra and args.workload == "c2" and args.n == n_def and args.seed == seed_def:
        del out, adj, off, status
        extra = measure_extra(lib, dev, cpu_check=not args.no_cpu)

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
        "config": {"workload": desc.format(n=args.n, seed=args.seed, total=n_cells),
                   "form": "batched" if clouds is not None else args.form, "n_points": n_cells,
                   "edges": e_total, "mean_faces": round(f_bar, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic,
                     "kernel": "cell construction: pg::" + " + pg::".join(
                         k for k in CELL_KERNELS if k != ("cell_group_exact_kernel" if clouds is not None
                                                          else "cell_group_exact_batch_kernel")),
                     "kernel_ms": round(kern_s * 1e3, 4),
                     "bytes_per_cell": round(b_cell, 2)},
        "cpu_baseline": cpu,
        "exact_cells": exact,
        "per_cloud_calls_ms": None if per_cloud_ms is None else round(per_cloud_ms, 3),
        "stats": instr.stats,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
