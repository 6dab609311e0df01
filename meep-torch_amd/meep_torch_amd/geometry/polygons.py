"""Simple polygons for the coverage rasterizer (the interface of
meep_torch/geometry/polygons.py: Polygon(vertices) with vertices (N, 2)
float64, any winding, implicitly closed; self-intersecting or degenerate
outlines raise ValueError, since the coverage of a non-simple outline is not
an area)."""
from dataclasses import dataclass

import numpy as np


def _cross(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def _proper_crossing(p, q, r, s):
    return ((_cross(r, s, p) > 0) != (_cross(r, s, q) > 0)) and ((_cross(p, q, r) > 0) != (_cross(p, q, s) > 0))


@dataclass(frozen=True)
class Polygon:
    vertices: np.ndarray

    def __post_init__(self):
        v = np.asarray(self.vertices, dtype=np.float64)
        if v.ndim != 2 or v.shape[1] != 2 or len(v) < 3:
            raise ValueError(f"need >=3 vertices of shape (N,2), got {v.shape}")
        n = len(v)
        for i in range(n):
            for j in range(i + 2, n):
                if i == 0 and j == n - 1:
                    continue          # adjacent through the closing edge
                if _proper_crossing(v[i], v[(i + 1) % n], v[j], v[(j + 1) % n]):
                    raise ValueError("polygon self-intersects")
        area = 0.5 * float(np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1]))
        if abs(area) < 1e-300:
            raise ValueError("degenerate polygon (zero area)")
        object.__setattr__(self, "vertices", v)

    @property
    def area(self):
        v = self.vertices
        return abs(0.5 * float(np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1])))
