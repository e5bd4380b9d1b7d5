"""MI355X-native 3D Voronoi and power diagrams with the API of paragram.

``voronoi_diagram(points)`` and ``power_diagram(points, weights)`` return a
``Diagram(adjacency, offsets, status)`` of int32 device tensors, as
paragram/_api.py:9-123 does.
"""
from ._api import Diagram, power_diagram, voronoi_diagram

__all__ = [
    "Diagram",
    "power_diagram",
    "voronoi_diagram",
]
