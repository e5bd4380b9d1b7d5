"""MI355X-native 3D Voronoi and power diagrams with the API of paragram.

``voronoi_diagram(points)`` and ``power_diagram(points, weights)`` return a
``Diagram(adjacency, offsets, status)`` of int32 device tensors, as
paragram/_api.py:9-123 does.  ``batched_voronoi_diagram`` and
``batched_power_diagram`` build many clouds in one launch (no counterpart in
the reference).
"""
from ._api import Diagram, batched_power_diagram, batched_voronoi_diagram, power_diagram, voronoi_diagram

__all__ = [
    "Diagram",
    "batched_power_diagram",
    "batched_voronoi_diagram",
    "power_diagram",
    "voronoi_diagram",
]
