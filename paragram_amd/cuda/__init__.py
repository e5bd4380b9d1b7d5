"""Engine entry points with the reference's names (paragram/cuda/__init__.py:1-12)."""
from .._backend import _C, _load_error

if _C is None:
    raise ImportError(
        "Paragram HIP extension is unavailable. A ROCm-enabled PyTorch install and the gfx950 library "
        f"libparagram_hip.so are required ({_load_error})."
    )

build_laguerre_voronoi_ultra = _C.build_laguerre_voronoi_ultra
build_voronoi_ultra = _C.build_voronoi_ultra

__all__ = [
    "build_laguerre_voronoi_ultra",
    "build_voronoi_ultra",
]
