"""Polygon geometry and the layout plugin against the reference.

The coverage rasterizer (accumulation buffer over per-cell edge pieces) is
an algorithm of its own; its output is pinned to the reference's coverage
(meep_torch/geometry/raster.py:37-67, golden `raster`, written by the
reference) to 1e-12 on a triangle, a rotated square, a concave L, a polygon
partly outside the grid, a clockwise pentagon and a 48-gon, and by what the
mathematics fixes: inside the grid the coverage sums to the polygon's area,
cells whose corners are all inside are exactly 1 and cells no part of the
outline reaches are exactly 0.  The painter's scene (TM eps, Kottke TE pair,
chi3) and the plugin's eps map are pinned to the reference's the same way."""
import numpy as np
import pytest
import torch

from conftest import golden

POLYS = {
    "triangle": [[0.11, 0.23], [1.37, 0.41], [0.52, 1.63]],
    "rot_square": [[0.8, 0.05], [1.55, 0.8], [0.8, 1.55], [0.05, 0.8]],
    "concave_l": [[0.2, 0.2], [1.4, 0.2], [1.4, 0.6], [0.6, 0.6], [0.6, 1.5], [0.2, 1.5]],
    "outside": [[-0.3, 0.9], [1.1, -0.25], [2.4, 1.3], [0.7, 2.2]],
    "cw_pentagon": [[0.9, 1.6], [1.5, 1.1], [1.3, 0.3], [0.5, 0.3], [0.3, 1.1]],
}


class _Grid:
    def __init__(self, nx, ny, dx):
        self.nx, self.ny, self.dx, self.device = nx, ny, dx, "cpu"


def _verts(G, name):
    return G["ngon"] if name == "ngon" else np.array(POLYS[name])


@pytest.mark.parametrize("name", list(POLYS) + ["ngon"])
def test_coverage_matches_reference(name):
    from meep_torch_amd.geometry import Polygon, coverage
    G = golden("raster")
    cov = coverage(Polygon(_verts(G, name)), _Grid(40, 36, 0.05))
    ref = torch.from_numpy(G[f"cov_{name}"])
    assert float((cov - ref).abs().max()) <= 1e-12


def _inside(p, v):
    x, y = p
    c = False
    for i in range(len(v)):
        (x0, y0), (x1, y1) = v[i], v[(i + 1) % len(v)]
        if (y0 > y) != (y1 > y) and x < x0 + (y - y0) * (x1 - x0) / (y1 - y0):
            c = not c
    return c


@pytest.mark.parametrize("name", ["triangle", "rot_square", "concave_l", "cw_pentagon"])
def test_coverage_area_and_exact_cells(name):
    from meep_torch_amd.geometry import Polygon, coverage_array
    v = np.array(POLYS[name])
    dx = 0.05
    cov = coverage_array(v, 40, 36, dx)
    # the area, up to the reference's snapping of cells within 1e-5 of full
    n_edge = int(((cov > 0) & (cov < 1)).sum())
    assert abs(cov.sum() * dx * dx - Polygon(v).area) <= 1e-5 * dx * dx * max(1, n_edge)
    for i in range(40):
        for j in range(36):
            corners = [_inside(((i + a) * dx, (j + b) * dx), v) for a in (0, 1) for b in (0, 1)]
            if all(corners):
                # all corners inside a polygon whose edges do not pass through the cell
                xs, ys = v[:, 0] / dx, v[:, 1] / dx
                if not np.any((xs > i) & (xs < i + 1) & (ys > j) & (ys < j + 1)):
                    assert cov[i, j] == 1.0, (i, j, cov[i, j])
    # cells far from the outline (more than one cell from any vertex box) are 0
    lo, hi = v.min(axis=0) / dx - 1, v.max(axis=0) / dx + 1
    far = np.ones_like(cov, dtype=bool)
    far[max(0, int(lo[0])):int(np.ceil(hi[0])) + 1, max(0, int(lo[1])):int(np.ceil(hi[1])) + 1] = False
    assert np.all(cov[far] == 0.0)


def test_coverage_of_a_rectangle_is_the_product_of_overlaps():
    """Closed form: an axis-aligned rectangle covers each cell by the product
    of its 1D overlaps (the facade's Block rule)."""
    from meep_torch_amd.geometry import coverage_array
    x0, y0, x1, y1, dx = 0.137, 0.221, 1.513, 0.968, 0.05
    cov = coverage_array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], 40, 36, dx)
    ox = np.clip(np.minimum((np.arange(40) + 1) * dx, x1) - np.maximum(np.arange(40) * dx, x0), 0, None) / dx
    oy = np.clip(np.minimum((np.arange(36) + 1) * dx, y1) - np.maximum(np.arange(36) * dx, y0), 0, None) / dx
    assert np.abs(cov - np.outer(ox, oy)).max() <= 1e-12


def test_self_intersecting_outline_raises():
    from meep_torch_amd.geometry import Polygon
    with pytest.raises(ValueError, match="self-intersects"):
        Polygon(np.array([[0, 0], [1, 1], [1, 0], [0, 1]], dtype=float))


def test_scene_matches_reference():
    from meep_torch_amd.geometry import Polygon, rasterize_scene
    G = golden("raster")
    ring = G["ngon"]
    items = [(Polygon(np.array(POLYS["concave_l"])), 12.0), (Polygon(ring), 4.0),
             (Polygon(np.array(POLYS["triangle"])), 2.25)]
    rm = rasterize_scene(items, _Grid(40, 36, 0.05), background=1.44, chi3_items=[(Polygon(ring), 0.02)])
    for mine, key in ((rm.eps_tm, "scene_eps_tm"), (rm.eps_te_xx, "scene_te_xx"), (rm.eps_te_yy, "scene_te_yy"),
                      (rm.chi3, "scene_chi3")):
        ref = torch.from_numpy(G[key])
        assert float((mine - ref).abs().max()) <= 1e-12 * float(ref.abs().max()), key


class _Port:
    def __init__(self, name, center, orientation):
        self.name, self.center, self.orientation = name, center, orientation


class _Component:
    """The two methods the plugin reads from a gdsfactory Component."""

    def __init__(self, polys, ports):
        self._p, self.ports = polys, ports

    def get_polygons_points(self, merge=False):
        assert merge is False
        return {1: [np.array(p) for p in self._p], 2: []}


def test_component_to_eps_matches_reference():
    from meep_torch_amd.plugins import component_to_eps
    G = golden("raster")
    polys = [[[0.0, 0.0], [2.0, 0.0], [2.0, 0.5], [0.0, 0.5]],
             [[2.0, 0.0], [3.1, 0.4], [3.3, 1.2], [2.9, 1.25], [2.6, 0.7], [2.0, 0.5]]]
    comp = _Component(polys, [_Port("o1", (0.0, 0.25), 180.0), _Port("o2", (3.1, 1.225), 90.0)])
    grid, eps, ports = component_to_eps(comp, res=20, pad=0.4, n_core=3.48, n_clad=1.44, device="cpu")
    assert (grid.nx, grid.ny) == (int(G["plugin_nx"]), int(G["plugin_ny"]))
    ref = torch.from_numpy(G["plugin_eps"])
    assert float((eps - ref).abs().max()) <= 1e-12 * float(ref.abs().max())
    # port centres -> nearest index from the padded origin (-0.4, -0.4),
    # Python's round-half-to-even as in the reference (1.625 um * 20 -> 32)
    assert ports == {"o1": (8, 13, 180.0), "o2": (70, 32, 90.0)}
    with pytest.raises(ValueError, match="no polygons"):
        component_to_eps(comp, res=20, pad=0.4, n_core=3.48, n_clad=1.44, layer=2, device="cpu")
