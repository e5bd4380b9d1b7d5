"""The learned-radius form: a small network predicts each cell's half-width,
the cell kernel starts from the box cut by that cube, and the exact traversal
certifies every cell (a cell the cube cuts is rebuilt from the box), so the
diagram is the plain method's, byte for byte; only the time changes.

No counterpart in the reference, whose traversal starts from the radius of a
kNN pre-clip (voronoi_ultra.cu:70-85, laguerre_ultra.cu:43-174).

    import paragram_amd.learned as pl
    model = pl.load()                        # the committed trained weights
    d = pl.power_diagram(points, weights, model=model)

Training (GPU, the library's own forward/backward/Adam kernels):
``python -m paragram_amd.learned --train`` fits the network to the half-widths
and face counts the plain method produces on seeded clouds drawn like the 10M
learned-power configuration, and writes ``radius_mlp.npz`` beside this file.
"""
from __future__ import annotations

import argparse
import ctypes
import dataclasses
import os
import time

import numpy as np
import torch

from . import _api, _engine, workloads

N_PARAMS = 5378
N_FEATURES = 16
MARGIN_QUANTILE = 0.98  # a cell starts inside its cube unless its residual exceeds this quantile


def interior(X: torch.Tensor) -> torch.Tensor:
    """Cells at least 8 neighbour distances inside the cloud's AABB (feature 9 = 1): the
    others reach towards the box and start from it (no cube)."""
    return X[:, 9] >= 1.0
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radius_mlp.npz")


@dataclasses.dataclass
class RadiusModel:
    params: np.ndarray      # float32 [5378]
    margin: float           # added to the predicted log(h / d_8): a high quantile of the residuals
    info: dict

    def device_params(self, dev) -> torch.Tensor:
        return torch.from_numpy(np.ascontiguousarray(self.params, dtype=np.float32)).to(dev)


def _lib():
    from ._backend import _C, _load_error

    if _C is None:
        raise ImportError(f"paragram_amd HIP engine unavailable: {_load_error}")
    return _C.lib


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    o = 0
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 2)):
        p[o:o + fan_in * fan_out] = rng.normal(0.0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


def features(lib, points: torch.Tensor, weights=None, leaf_size: int = -1):
    """``(X float32 [N, 16], log_d8 float32 [N])`` on points' device (original order)."""
    n = int(points.shape[0])
    dev = points.device
    X = torch.empty((n, N_FEATURES), dtype=torch.float32, device=dev)
    logd8 = torch.empty(n, dtype=torch.float32, device=dev)
    if n == 0:
        return X, logd8
    with torch.cuda.device(dev):
        sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        ws_bytes = lib.u64("pg_index_workspace_bytes", n)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        lib.call("pg_build_index", n, _engine._p(points), _engine._p(weights), int(leaf_size), _engine._p(ws),
                 ws_bytes, sp)
        ids = torch.empty((n, 8), dtype=torch.int32, device=dev)
        d2 = torch.empty((n, 8), dtype=torch.float32, device=dev)
        lib.call("pg_knn", n, _engine._p(ws), 8, _engine._p(ids), _engine._p(d2), sp)
        lib.call("pg_mlp_features", n, _engine._p(ws), _engine._p(ids), _engine._p(d2), _engine._p(X),
                 _engine._p(logd8), sp)
    return X, logd8


def forward(lib, X: torch.Tensor, params: torch.Tensor) -> torch.Tensor:
    """Network outputs ``[N, 2]`` (matrix-core kernel)."""
    n = int(X.shape[0])
    Y = torch.empty((n, 2), dtype=torch.float32, device=X.device)
    if n:
        sp = ctypes.c_void_p(torch.cuda.current_stream(X.device).cuda_stream)
        lib.call("pg_mlp_forward", n, _engine._p(X), _engine._p(params), _engine._p(Y), sp)
    return Y


def targets(lib, points: torch.Tensor, weights, logd8: torch.Tensor):
    """Training targets from the plain method: ``(T [N, 2], mask [N])`` with
    T = (log(h / d_8), faces / 16); cells that are empty or not built cleanly
    are masked out."""
    n = int(points.shape[0])
    ext = torch.zeros(n, dtype=torch.float32, device=points.device)
    adj, off, st = _engine.build_diagram(lib, points, weights, extent=ext)
    faces = (off[1:] - off[:-1]).to(torch.float32)
    ok = (ext > 0) & (st == 0)
    T = torch.stack([torch.log(torch.clamp(ext, min=1e-30)) - logd8, faces / 16.0], dim=1)
    return T.contiguous(), ok.to(torch.float32)


class Trainer:
    """Adam on the GPU through pg_mlp_train_step."""

    def __init__(self, lib, params: np.ndarray, dev, batch: int):
        self.lib = lib
        self.dev = dev
        self.P = torch.from_numpy(params.astype(np.float32)).to(dev)
        self.m = torch.zeros(N_PARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(N_PARAMS, dtype=torch.float32, device=dev)
        self.grad = torch.zeros(N_PARAMS + 1, dtype=torch.float32, device=dev)
        self.part = torch.empty(lib.u64("pg_mlp_train_scratch_elems", batch), dtype=torch.float32, device=dev)
        self.t = 0

    def step(self, X, T, mask, count: float, lr: float):
        """One step on the rows given (lr = 0: gradient only).  Returns the mean loss tensor."""
        if lr > 0:
            self.t += 1
        sp = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        p = _engine._p
        self.lib.call("pg_mlp_train_step", int(X.shape[0]), p(X), p(T), p(mask), ctypes.c_float(count), p(self.P),
                      p(self.grad), p(self.m), p(self.v), p(self.part), ctypes.c_float(lr), max(self.t, 1), sp)
        return self.grad[N_PARAMS]


def train(lib, clouds, dev, epochs: int = 3, batch: int = 8192, lr: float = 2e-3, seed: int = 0,
          rows_per_cloud: int = 200_000, log=print):
    """Fits the network to the plain method's cells on ``clouds`` (list of (points, weights)
    numpy arrays); returns a RadiusModel."""
    Xs, Ts, Ms = [], [], []
    for pts, w in clouds:
        P = torch.from_numpy(pts).to(dev)
        W = None if w is None else torch.from_numpy(w).to(dev).reshape(-1, 1)
        X, logd8 = features(lib, P, W)
        T, mask = targets(lib, P, W, logd8)
        sel = torch.randperm(P.shape[0], device=dev, generator=torch.Generator(device=dev).manual_seed(seed))
        sel = sel[:rows_per_cloud]
        Xs.append(X[sel])
        Ts.append(T[sel])
        Ms.append(mask[sel])
        seed += 1
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    M = torch.cat(Ms)
    n = X.shape[0]
    n_val = min(n // 10, 200_000)
    Xv, Tv, Mv = X[:n_val], T[:n_val], M[:n_val]
    X, T, M = X[n_val:], T[n_val:], M[n_val:]
    tr = Trainer(lib, init_params(seed), dev, batch)
    g = torch.Generator(device=dev).manual_seed(seed)
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = torch.randperm(X.shape[0], device=dev, generator=g)
        rate = lr * (0.3 ** ep)
        for b in range(0, X.shape[0] - batch + 1, batch):
            idx = order[b:b + batch]
            xb, tb, mb = X[idx].contiguous(), T[idx].contiguous(), M[idx].contiguous()
            tr.step(xb, tb, mb, float(mb.sum().item()), rate)
        log(f"epoch {ep}: train loss {float(tr.grad[N_PARAMS]):.4f} ({time.perf_counter() - t0:.1f}s)")
    Y = forward(lib, Xv, tr.P)
    ok = (Mv > 0) & interior(Xv)
    res = (Tv[:, 0] - Y[:, 0])[ok]
    margin = float(torch.quantile(res, MARGIN_QUANTILE).item())
    rmse = float(torch.sqrt((res ** 2).mean()).item())
    log(f"validation (interior cells): log-radius rmse {rmse:.4f}, {MARGIN_QUANTILE:.0%} residual quantile "
        f"{margin:.4f}")
    info = {"rows": int(n), "epochs": epochs, "val_rmse_log_h": rmse, "margin_quantile": MARGIN_QUANTILE}
    return RadiusModel(tr.P.cpu().numpy(), margin, info)


def predict_on_index(lib, model: RadiusModel, ws: torch.Tensor, n: int, dev, sp, n_clouds: int = 0) -> torch.Tensor:
    """Predicted half-widths from an index already built (pg_build_index, or
    pg_build_batch_index with n_clouds > 0: neighbours and hull features per
    cloud): kNN, features and the network on that workspace, no second index."""
    p = _engine._p
    ids = torch.empty((n, 8), dtype=torch.int32, device=dev)
    d2 = torch.empty((n, 8), dtype=torch.float32, device=dev)
    X = torch.empty((n, N_FEATURES), dtype=torch.float32, device=dev)
    logd8 = torch.empty(n, dtype=torch.float32, device=dev)
    Y = torch.empty((n, 2), dtype=torch.float32, device=dev)
    if n_clouds > 0:
        lib.call("pg_knn_batch", n, p(ws), n_clouds, 8, p(ids), p(d2), sp)
        lib.call("pg_mlp_features_batch", n, p(ws), n_clouds, p(ids), p(d2), p(X), p(logd8), sp)
    else:
        lib.call("pg_knn", n, p(ws), 8, p(ids), p(d2), sp)
        lib.call("pg_mlp_features", n, p(ws), p(ids), p(d2), p(X), p(logd8), sp)
    lib.call("pg_mlp_forward", n, p(X), p(model.device_params(dev)), p(Y), sp)
    h = torch.exp(Y[:, 0] + logd8 + model.margin)
    return torch.where(interior(X), h, torch.zeros_like(h)).contiguous()


def predict_half_width(lib, model: RadiusModel, points: torch.Tensor, weights=None) -> torch.Tensor:
    """Predicted half-width of every cell (float32 [N], original order)."""
    X, logd8 = features(lib, points, weights)
    Y = forward(lib, X, model.device_params(points.device))
    h = torch.exp(Y[:, 0] + logd8 + model.margin)
    return torch.where(interior(X), h, torch.zeros_like(h)).contiguous()


def save(model: RadiusModel, path: str = MODEL_PATH) -> None:
    np.savez(path, params=model.params.astype(np.float32), margin=np.float32(model.margin),
             info=np.array(repr(model.info)))


def load(path: str = MODEL_PATH) -> RadiusModel:
    with np.load(path, allow_pickle=False) as f:
        return RadiusModel(f["params"].astype(np.float32), float(f["margin"]), {"source": path})


def power_diagram(points: torch.Tensor, weights: torch.Tensor, *, model: RadiusModel | None = None,
                  leaf_size: int = -1) -> _api.Diagram:
    """``paragram_amd.power_diagram`` started from the learned radius (same result)."""
    points = _api._validate_points(points)
    weights = _api._validate_weights(weights, points.shape[0], points.device)
    lib = _lib()
    model = model or load()
    a, o, s = _engine.build_diagram(lib, points, weights, 8, False, None, None, leaf_size, radius_model=model)
    return _api.Diagram(adjacency=a, offsets=o, status=s)


def voronoi_diagram(points: torch.Tensor, *, model: RadiusModel | None = None, leaf_size: int = -1) -> _api.Diagram:
    """``paragram_amd.voronoi_diagram`` started from the learned radius (same result)."""
    points = _api._validate_points(points)
    lib = _lib()
    model = model or load()
    a, o, s = _engine.build_diagram(lib, points, None, 8, False, None, None, leaf_size, radius_model=model)
    return _api.Diagram(adjacency=a, offsets=o, status=s)


def batched_voronoi_diagram(points: torch.Tensor, cloud_offsets, *, model: RadiusModel | None = None,
                            leaf_size: int = -1) -> list:
    """``paragram_amd.batched_voronoi_diagram`` with each cell started from the
    learned radius (neighbours and features per cloud; same result)."""
    points = _api._validate_points(points)
    off = _api._validate_cloud_offsets(cloud_offsets, points.shape[0])
    lib = _lib()
    model = model or load()
    a, o, s = _engine.build_diagram(lib, points, None, 8, False, None, None, leaf_size, clouds=off,
                                    radius_model=model)
    return [_api.Diagram(adjacency=x, offsets=y, status=z) for x, y, z in _engine.split_batch(a, o, s, off, lib=lib)]


def batched_power_diagram(points: torch.Tensor, weights: torch.Tensor, cloud_offsets, *,
                          model: RadiusModel | None = None, leaf_size: int = -1) -> list:
    """``paragram_amd.batched_power_diagram`` with each cell started from the
    learned radius (same result)."""
    points = _api._validate_points(points)
    weights = _api._validate_weights(weights, points.shape[0], points.device)
    off = _api._validate_cloud_offsets(cloud_offsets, points.shape[0])
    lib = _lib()
    model = model or load()
    a, o, s = _engine.build_diagram(lib, points, weights, 8, False, None, None, leaf_size, clouds=off,
                                    radius_model=model)
    return [_api.Diagram(adjacency=x, offsets=y, status=z) for x, y, z in _engine.split_batch(a, o, s, off, lib=lib)]


def training_clouds(count: int = 50, n: int = 1_000_000, seed: int = 1000):
    """Seeded clouds drawn like the learned-power configuration (50% uniform +
    50% from 256 Gaussians, weights U[0, 5e-6]), one seed each."""
    for k in range(count):
        yield workloads.learned_power(n, seed + k)


def main() -> int:
    ap = argparse.ArgumentParser(description="train the learned-radius network on the GPU")
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--clouds", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--out", default=MODEL_PATH)
    args = ap.parse_args()
    if not args.train:
        ap.print_help()
        return 2
    dev = torch.device("cuda:0")
    lib = _lib()
    t0 = time.perf_counter()
    model = train(lib, list(training_clouds(args.clouds, args.n)), dev, epochs=args.epochs,
                  log=lambda m: print(m, flush=True))
    model.info.update({"clouds": args.clouds, "points_per_cloud": args.n, "seeds": [1000, 1000 + args.clouds - 1],
                       "seconds": time.perf_counter() - t0})
    save(model, args.out)
    print(f"wrote {args.out} ({model.info})", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
