"""Loads the HIP engine and exposes it as ``_C`` with the reference's two callables.

Mirrors the seam of the reference (paragram/_backend.py:138-195): ``_C`` is an
object with ``build_voronoi_ultra`` and ``build_laguerre_voronoi_ultra``, or
``None`` when the engine cannot be loaded, in which case importing
``paragram_amd.cuda`` raises ``ImportError`` (paragram/cuda/__init__.py:3-4).
The engine is the in-tree ``libparagram_hip.so`` (built by
``paragram_amd._build``); there is no CPU fallback.
"""
from __future__ import annotations

import os
from typing import Optional

import torch

from . import _engine

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


def _load() -> Optional[_HipBackend]:
    global _load_error
    if not os.path.exists(_engine.LIB_PATH):
        _load_error = f"{_engine.LIB_PATH} is missing (build it with paragram_amd._build.build())"
        return None
    try:
        return _HipBackend(_engine.HipLibrary(_engine.LIB_PATH))
    except OSError as exc:
        _load_error = str(exc)
        return None


_C = _load()
