"""Loads the HIP engine and exposes it as ``_C`` with the reference's two callables.

Mirrors the seam of the reference (paragram/_backend.py:138-195): ``_C`` is an
object with ``build_voronoi_ultra`` and ``build_laguerre_voronoi_ultra``, or
``None`` when the engine cannot be loaded, in which case importing
``paragram_amd.cuda`` raises ``ImportError`` (paragram/cuda/__init__.py:3-4).
The engine is the in-tree ``libparagram_hip.so``.  As the reference compiles
its extension on first import when no build exists (_backend.py:147-190),
a missing library is built here on first import with hipcc for gfx950
(``paragram_amd._build``: ``MAX_JOBS`` (default 10), ``FAST_COMPILE``,
``DEBUG_CUDA``, ``PROFILE_CUDA``, ``VERBOSE``), unless
``PARAGRAM_NO_BUILD=1``; without hipcc the engine is unavailable.  There is
no CPU fallback.
"""
from __future__ import annotations

import os
import time
from typing import Optional

import torch

from . import _build, _engine

_load_error: Optional[str] = None


class _HipBackend:
    """Drop-in for the reference's pybind module (paragram/csrc/ext.cpp:10-29)."""

    def __init__(self, lib: _engine.HipLibrary):
        self.lib = lib

    def build_voronoi_ultra(self, points: torch.Tensor, knn: int = 8, debug: bool = False,
                            initial_guesses: Optional[torch.Tensor] = None,
                            initial_guesses_offsets: Optional[torch.Tensor] = None, leaf_size: int = -1):
        return _engine.build_diagram(self.lib, points, None, knn, debug, initial_guesses,
                                     initial_guesses_offsets, leaf_size)

    def build_laguerre_voronoi_ultra(self, points: torch.Tensor, weights: torch.Tensor, knn: int = 8,
                                     debug: bool = False, initial_guesses: Optional[torch.Tensor] = None,
                                     initial_guesses_offsets: Optional[torch.Tensor] = None, leaf_size: int = -1):
        return _engine.build_diagram(self.lib, points, weights, knn, debug, initial_guesses,
                                     initial_guesses_offsets, leaf_size)


def _build_missing(lib_path: str) -> bool:
    """Builds the library when it is missing; False (and the reason in
    _load_error) when it cannot be built here."""
    global _load_error
    if os.environ.get("PARAGRAM_NO_BUILD") == "1":
        _load_error = f"{lib_path} is missing and PARAGRAM_NO_BUILD=1"
        return False
    try:
        _build._hipcc()
    except RuntimeError as exc:
        _load_error = f"{lib_path} is missing and cannot be built: {exc}"
        print(f"paragram_amd: {_load_error}; the HIP engine is unavailable.", flush=True)
        return False
    jobs = os.environ.get("MAX_JOBS") or "10"
    print(f"paragram_amd: building the HIP library for gfx950 with MAX_JOBS={jobs}", flush=True)
    t0 = time.time()
    prev = os.environ.get("MAX_JOBS")
    os.environ["MAX_JOBS"] = jobs
    try:
        _build.build(lib_path=lib_path)
    except RuntimeError as exc:
        _load_error = f"building {lib_path} failed: {exc}"
        return False
    finally:
        if prev is None:
            os.environ.pop("MAX_JOBS", None)
    print(f"paragram_amd: HIP library ready in {time.time() - t0:.2f} seconds.", flush=True)
    return True


def _load(lib_path: str = _engine.LIB_PATH) -> Optional[_HipBackend]:
    global _load_error
    if not os.path.exists(lib_path) and not _build_missing(lib_path):
        return None
    try:
        return _HipBackend(_engine.HipLibrary(lib_path))
    except OSError as exc:
        _load_error = str(exc)
        return None


_C = _load()
