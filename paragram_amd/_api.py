"""Public API: ``Diagram``, ``voronoi_diagram``, ``power_diagram``.

Signatures, keyword defaults, the order of the checks, the exception types
and the messages are those of paragram/_api.py:9-123, so code written against
the reference sees the same behaviour.  Points on a ROCm device report
``is_cuda`` True under PyTorch-ROCm.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch


@dataclass(frozen=True)
class Diagram:
    """CSR neighbour graph of the cells (paragram/_api.py:9-13).

    adjacency: int32 [E], neighbour ids, each row sorted ascending.
    offsets:   int32 [N + 1], row i is adjacency[offsets[i]:offsets[i + 1]].
    status:    int32 [N], 0 success, 3 empty (hidden) cell.
    """

    adjacency: torch.Tensor
    offsets: torch.Tensor
    status: torch.Tensor


def _require_device_tensor(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on a CUDA device")


def _validate_points(points: torch.Tensor) -> torch.Tensor:
    # paragram/_api.py:16-25
    _require_device_tensor(points, "points")
    if points.dtype != torch.float32:
        raise ValueError("points must have dtype torch.float32")
    if points.ndim != 2 or points.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    return points.contiguous()


def _validate_weights(weights: torch.Tensor, num_points: int, device: torch.device) -> torch.Tensor:
    # paragram/_api.py:28-41
    _require_device_tensor(weights, "weights")
    if weights.device != device:
        raise ValueError("weights must be on the same device as points")
    if weights.dtype != torch.float32:
        raise ValueError("weights must have dtype torch.float32")
    w = weights[:, None] if weights.ndim == 1 else weights
    if w.ndim != 2 or tuple(w.shape) != (num_points, 1):
        raise ValueError("weights must have shape (N,) or (N, 1)")
    return w.contiguous()


def _validate_initial_guesses(
    initial_guesses: Optional[torch.Tensor],
    initial_guesses_offsets: Optional[torch.Tensor],
    num_points: int,
    device: torch.device,
) -> tuple[Optional[torch.Tensor], Optional[torch.Tensor]]:
    # paragram/_api.py:44-62
    if initial_guesses is None and initial_guesses_offsets is None:
        return None, None
    if initial_guesses is None or initial_guesses_offsets is None:
        raise ValueError("initial_guesses and initial_guesses_offsets must be provided together")
    pair = (initial_guesses, initial_guesses_offsets)
    if not all(t.is_cuda for t in pair):
        raise ValueError("initial guess tensors must be on a CUDA device")
    if any(t.device != device for t in pair):
        raise ValueError("initial guess tensors must be on the same device as points")
    if any(t.dtype != torch.int32 for t in pair):
        raise ValueError("initial guess tensors must have dtype torch.int32")
    if initial_guesses_offsets.ndim != 1 or initial_guesses_offsets.shape[0] != num_points + 1:
        raise ValueError("initial_guesses_offsets must have shape (N + 1,)")
    return initial_guesses.contiguous(), initial_guesses_offsets.contiguous()


def power_diagram(
    points: torch.Tensor,
    weights: torch.Tensor,
    *,
    knn: int = 8,
    debug: bool = False,
    initial_guesses: Optional[torch.Tensor] = None,
    initial_guesses_offsets: Optional[torch.Tensor] = None,
    leaf_size: int = -1,
) -> Diagram:
    """Power (Laguerre) diagram of weighted points clipped to the padded box.

    Cell i is {x : |x - p_i|^2 - w_i <= |x - p_j|^2 - w_j for all j} within
    [min - 1, max + 1]; hidden points get status 3 and an empty row.
    """
    points = _validate_points(points)
    weights = _validate_weights(weights, points.shape[0], points.device)
    initial_guesses, initial_guesses_offsets = _validate_initial_guesses(
        initial_guesses, initial_guesses_offsets, points.shape[0], points.device
    )
    from .cuda import build_laguerre_voronoi_ultra

    adjacency, offsets, status = build_laguerre_voronoi_ultra(
        points,
        weights,
        knn=knn,
        debug=debug,
        initial_guesses=initial_guesses,
        initial_guesses_offsets=initial_guesses_offsets,
        leaf_size=leaf_size,
    )
    return Diagram(adjacency=adjacency, offsets=offsets, status=status)


def voronoi_diagram(
    points: torch.Tensor,
    *,
    knn: int = 8,
    debug: bool = False,
    initial_guesses: Optional[torch.Tensor] = None,
    initial_guesses_offsets: Optional[torch.Tensor] = None,
    leaf_size: int = -1,
) -> Diagram:
    """Voronoi diagram of points clipped to the padded box [min - 1, max + 1]."""
    points = _validate_points(points)
    initial_guesses, initial_guesses_offsets = _validate_initial_guesses(
        initial_guesses, initial_guesses_offsets, points.shape[0], points.device
    )
    from .cuda import build_voronoi_ultra

    adjacency, offsets, status = build_voronoi_ultra(
        points,
        knn=knn,
        debug=debug,
        initial_guesses=initial_guesses,
        initial_guesses_offsets=initial_guesses_offsets,
        leaf_size=leaf_size,
    )
    return Diagram(adjacency=adjacency, offsets=offsets, status=status)


# ---------------------------------------------------------------- batches
# No counterpart in the reference (one cloud per call there): many clouds in
# one persistent launch, each cloud's result equal to its single call.

def _validate_cloud_offsets(cloud_offsets, num_points: int) -> "np.ndarray":
    import numpy as np

    if isinstance(cloud_offsets, torch.Tensor):
        cloud_offsets = cloud_offsets.detach().cpu().numpy()
    off = np.asarray(cloud_offsets)
    if off.ndim != 1 or off.shape[0] < 2 or not np.issubdtype(off.dtype, np.integer):
        raise ValueError("cloud_offsets must be a 1-D integer sequence of length (B + 1,) with B >= 1")
    off = off.astype(np.int64)
    if off.shape[0] - 1 > 4096:
        raise ValueError("at most 4096 clouds per batch")
    if off[0] != 0 or off[-1] != num_points or (np.diff(off) < 0).any():
        raise ValueError("cloud_offsets must start at 0, end at N and be non-decreasing")
    return off


def _batch(points, weights, cloud_offsets, knn, debug, leaf_size):
    from ._backend import _C, _load_error
    from ._engine import build_diagram, split_batch

    if _C is None:
        raise ImportError(f"paragram_amd HIP engine unavailable: {_load_error}")
    off = _validate_cloud_offsets(cloud_offsets, points.shape[0])
    adjacency, offsets, status = build_diagram(_C.lib, points, weights, knn, debug, None, None, leaf_size, clouds=off)
    return [Diagram(adjacency=a, offsets=o, status=s) for a, o, s in split_batch(adjacency, offsets, status, off)]


def batched_voronoi_diagram(points: torch.Tensor, cloud_offsets, *, knn: int = 8, debug: bool = False,
                            leaf_size: int = -1) -> list:
    """Voronoi diagrams of a batch of clouds in one launch.

    ``points`` (N, 3) float32 holds the clouds back to back; cloud c is rows
    ``cloud_offsets[c]`` .. ``cloud_offsets[c + 1]``.  Returns one ``Diagram``
    per cloud, equal to ``voronoi_diagram`` of that cloud alone (cloud-local
    ids, each cloud clipped to its own padded box).
    """
    points = _validate_points(points)
    return _batch(points, None, cloud_offsets, knn, debug, leaf_size)


def batched_power_diagram(points: torch.Tensor, weights: torch.Tensor, cloud_offsets, *, knn: int = 8,
                          debug: bool = False, leaf_size: int = -1) -> list:
    """Power diagrams of a batch of weighted clouds (see batched_voronoi_diagram)."""
    points = _validate_points(points)
    weights = _validate_weights(weights, points.shape[0], points.device)
    return _batch(points, weights, cloud_offsets, knn, debug, leaf_size)
