"""Layout-component plugin (the interface of
meep_torch/plugins/gdsfactory.py:19-42): a gdsfactory-style Component ->
(grid, eps, port index map) for the 2D engine.

The component is duck-typed, so gdsfactory itself is not needed: anything
with ``get_polygons_points(merge=False)`` (a dict layer -> list of (N, 2)
vertex arrays, in um) and ``ports`` (objects with ``name``, ``center``,
``orientation``) works; ``copy()`` is used when present.  Effective-index
2D: the polygons of `layer` are painted with n_core^2 over n_clad^2 by exact
area coverage; the padded bounding box of the polygons is shifted to start
at (0, 0); port centres map to the nearest grid index."""
import numpy as np

from ..core.grid import Grid
from ..geometry.polygons import Polygon
from ..geometry.scene import rasterize_scene


def component_to_eps(component, *, res, pad, n_core, n_clad, layer=1, dtype=None, device=None):
    comp = component.copy() if hasattr(component, "copy") else component
    # unmerged pieces: merged rings come back as keyhole outlines that touch
    # themselves, which the rasterizer rejects as non-simple
    polys = [np.asarray(p, dtype=np.float64) for p in comp.get_polygons_points(merge=False)[layer]]
    if not polys:
        raise ValueError(f"component has no polygons on layer {layer!r}")
    allv = np.concatenate(polys, axis=0)
    lo = allv.min(axis=0) - pad
    span = allv.max(axis=0) + pad - lo
    kw = {}
    if dtype is not None:
        kw["dtype"] = dtype
    if device is not None:
        kw["device"] = device
    grid = Grid(nx=int(round(span[0] * res)), ny=int(round(span[1] * res)), dx=1.0 / res, **kw)
    items = [(Polygon(p - lo), n_core ** 2) for p in polys]
    rm = rasterize_scene(items, grid, background=n_clad ** 2)
    ports = {p.name: (int(round((p.center[0] - lo[0]) * res)), int(round((p.center[1] - lo[1]) * res)),
                      p.orientation) for p in comp.ports}
    return grid, rm.eps_tm, ports
