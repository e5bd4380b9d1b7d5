"""Autotuner for the launch choices that change speed but never the diagram:
the LBVH leaf size, the number of nearest neighbours clipped first, and the
persistent waves per compute unit of the cell launch.

``tune(points, weights)`` times every candidate on a bounded prefix of the
caller's own data (2 calls each after one warm-up) and returns the fastest;
results are cached per (size bucket, weighted, batched) key for the process
and, if ``PARAGRAM_AUTOTUNE_CACHE`` names a file, in that JSON file.  The
tuned entry points (``voronoi_diagram``, ``power_diagram``,
``batched_voronoi_diagram``) return exactly what the untuned ones do.

No counterpart in the reference (its leaf size and knn are fixed defaults,
voronoi_ultra.cu:175-182).
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np
import torch

from . import _api, _engine

LEAF_SIZES = (8, 16, 32)
PRECLIPS = (8, 16)
WAVES_PER_CU = (12, 16, 24)  # 12 = what tier 0 keeps resident (3 per SIMD); more queue behind them
SAMPLE = 200_000
_cache: dict = {}


def _key(n: int, weighted: bool, batched: bool) -> str:
    return f"n2^{max(int(math.log2(max(n, 1))), 0)}_{'power' if weighted else 'voronoi'}_{'batch' if batched else 'one'}"


def _load_file():
    path = os.environ.get("PARAGRAM_AUTOTUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            _cache.update(json.load(f))


def _save_file():
    path = os.environ.get("PARAGRAM_AUTOTUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump(_cache, f, indent=1)


def tune(points: torch.Tensor, weights: torch.Tensor | None = None, clouds=None, sample: int = SAMPLE,
         force: bool = False) -> dict:
    """The fastest {leaf_size, preclip, waves_per_cu} for data like ``points`` (timed on up to
    ``sample`` of them): leaf size and pre-clip over their grid at the default waves, then the
    waves per compute unit at the best of those."""
    from ._backend import _C

    n = int(points.shape[0])
    key = _key(n, weights is not None, clouds is not None)
    if not _cache:
        _load_file()
    if key in _cache and not force:
        return dict(_cache[key])
    m = min(n, sample)
    pts = points[:m].contiguous()
    w = None if weights is None else weights[:m].contiguous()
    cl = None
    if clouds is not None:  # the clouds that fit in the sample, whole
        off = np.asarray(clouds, dtype=np.int64)
        keep = off[off <= m]
        cl = keep if keep.shape[0] >= 2 else np.array([0, m], np.int64)
        m = int(cl[-1])
        pts, w = pts[:m], None if w is None else w[:m]
    best, times = None, {}

    def timed(cfg):
        kw = dict(clouds=cl, preclip=cfg["preclip"], waves_per_cu=cfg["waves_per_cu"])
        _engine.build_diagram(_C.lib, pts, w, 8, False, None, None, cfg["leaf_size"], **kw)
        torch.cuda.synchronize(points.device)
        t0 = time.perf_counter()
        for _ in range(2):
            _engine.build_diagram(_C.lib, pts, w, 8, False, None, None, cfg["leaf_size"], **kw)
        torch.cuda.synchronize(points.device)
        t = (time.perf_counter() - t0) / 2
        times[f"{cfg['leaf_size']}/{cfg['preclip']}/{cfg['waves_per_cu']}"] = round(t * 1e3, 3)
        return t

    for leaf in LEAF_SIZES:
        for pre in PRECLIPS:
            cfg = {"leaf_size": leaf, "preclip": pre, "waves_per_cu": 16}
            t = timed(cfg)
            if best is None or t < best[0]:
                best = (t, cfg)
    for wpc in WAVES_PER_CU:
        if wpc == 16:
            continue
        cfg = dict(best[1], waves_per_cu=wpc)
        t = timed(cfg)
        if t < best[0]:
            best = (t, cfg)
    out = dict(best[1], sample=m, ms=times)
    _cache[key] = out
    _save_file()
    return dict(out)


def voronoi_diagram(points: torch.Tensor) -> _api.Diagram:
    """``paragram_amd.voronoi_diagram`` with tuned launch choices (same result)."""
    from ._backend import _C

    points = _api._validate_points(points)
    cfg = tune(points)
    a, o, s = _engine.build_diagram(_C.lib, points, None, 8, False, None, None, cfg["leaf_size"],
                                    preclip=cfg["preclip"], waves_per_cu=cfg.get("waves_per_cu"))
    return _api.Diagram(adjacency=a, offsets=o, status=s)


def power_diagram(points: torch.Tensor, weights: torch.Tensor) -> _api.Diagram:
    """``paragram_amd.power_diagram`` with tuned launch choices (same result)."""
    from ._backend import _C

    points = _api._validate_points(points)
    weights = _api._validate_weights(weights, points.shape[0], points.device)
    cfg = tune(points, weights)
    a, o, s = _engine.build_diagram(_C.lib, points, weights, 8, False, None, None, cfg["leaf_size"],
                                    preclip=cfg["preclip"], waves_per_cu=cfg.get("waves_per_cu"))
    return _api.Diagram(adjacency=a, offsets=o, status=s)


def batched_voronoi_diagram(points: torch.Tensor, cloud_offsets) -> list:
    """``paragram_amd.batched_voronoi_diagram`` with tuned launch choices (same result)."""
    from ._backend import _C

    points = _api._validate_points(points)
    off = _api._validate_cloud_offsets(cloud_offsets, points.shape[0])
    cfg = tune(points, None, off)
    a, o, s = _engine.build_diagram(_C.lib, points, None, 8, False, None, None, cfg["leaf_size"], clouds=off,
                                    preclip=cfg["preclip"], waves_per_cu=cfg.get("waves_per_cu"))
    return [_api.Diagram(adjacency=x, offsets=y, status=z) for x, y, z in _engine.split_batch(a, o, s, off)]
