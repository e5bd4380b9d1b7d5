"""Memory-admission planner for the per-cell row pool.

Same environment semantics as the reference's plan_convex_cell_chunks
(paragram/csrc/laguerre/chunking.h:16-56):

* ``PARAGRAM_NUM_CHUNKS=n`` (n > 0) forces ``chunk_blocks = ceil(total / n)``;
* otherwise ``PARAGRAM_MEM_FRAC`` (used only if 0.05 < value < 0.95, default
  0.5) times the free device memory is the budget, and when the full slab
  exceeds it, ``chunks = ceil(full / budget)``;
* ``chunk_blocks`` is clamped to [1, total] and the chunk count recomputed, so
  the result can be smaller than the forced count.

The admission unit here is a block of 64 Morton-consecutive cells (one
wavefront) and the chunked resource is the row pool of those cells (32
entries per cell by default), the larger of the two per-cell buffers (the
cell scratch takes 32 bytes per cell).
"""
from __future__ import annotations

import os
from dataclasses import dataclass


@dataclass(frozen=True)
class ChunkPlan:
    chunk_blocks: int
    chunk_threads: int
    num_chunks: int


def _atoi(s: str) -> int:
    """C atoi: leading whitespace, optional sign, digits; 0 if none."""
    s = s.lstrip()
    sign, i = 1, 0
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def _atof(s: str) -> float:
    """C atof on the longest numeric prefix; 0.0 if none."""
    s = s.strip()
    best = 0.0
    for end in range(len(s), 0, -1):
        try:
            best = float(s[:end])
            break
        except ValueError:
            continue
    return best


def plan_chunks(total_blocks: int, block_stride: int, full_bytes: int, free_bytes: int,
                env: dict | None = None) -> ChunkPlan:
    env = os.environ if env is None else env
    chunk_blocks = total_blocks
    force = env.get("PARAGRAM_NUM_CHUNKS")
    if force is not None and _atoi(force) > 0:
        n = _atoi(force)
        chunk_blocks = (total_blocks + n - 1) // n
    else:
        frac = 0.5
        fenv = env.get("PARAGRAM_MEM_FRAC")
        if fenv is not None:
            parsed = _atof(fenv)
            if 0.05 < parsed < 0.95:
                frac = parsed
        budget = int(frac * float(free_bytes))
        if budget > 0 and full_bytes > budget:
            n = (full_bytes + budget - 1) // budget
            chunk_blocks = (total_blocks + n - 1) // n
    chunk_blocks = max(1, min(chunk_blocks, total_blocks))
    num_chunks = (total_blocks + chunk_blocks - 1) // chunk_blocks
    return ChunkPlan(chunk_blocks, chunk_blocks * block_stride, num_chunks)
