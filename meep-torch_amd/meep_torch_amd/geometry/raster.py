"""Exact area coverage of polygons on the engine grid (the interface of
meep_torch/geometry/raster.py:37-67: coverage(poly, grid) -> (nx, ny)
float64 fraction of every cell the polygon covers, snapped to exact 0 / 1
within 1e-9).

Method: an accumulation buffer.  Every polygon edge is cut where it crosses
integer x or y (grid units), so each piece lies in one cell (i, j).  A piece
of signed height h (its y extent, signed by the edge direction) covers, in
its own cell, the area to its right, h * (i + 1 - x_mid), and the full h in
every cell further right along the row.  Adding h * (i + 1 - x_mid) to
acc[i, j] and h * (x_mid - i) to acc[i + 1, j] and summing acc along x gives
the winding-weighted area of every cell; a simple polygon's orientation
fixes the sign.  Pieces left of the grid add their full h to column 0; rows
and columns beyond it are dropped.  Nearly full and empty cells are then
snapped as the reference snaps them (see snap).  The edges are processed together with
numpy, in O(pieces) -- no per-row polygon clipping."""
import numpy as np
import torch


def _pieces(v):
    """Cut every edge of the closed polygon v (N, 2), in grid units, at its
    integer x and y crossings: returns (x_mid, y_mid, h) of every piece."""
    a = v
    b = np.roll(v, -1, axis=0)
    xs, ys, hs = [], [], []
    for (x0, y0), (x1, y1) in zip(a, b):
        if y0 == y1:
            continue                      # horizontal: no area to its right
        t = [0.0, 1.0]
        for c0, c1 in ((x0, x1), (y0, y1)):
            if c0 != c1:
                lo, hi = min(c0, c1), max(c0, c1)
                k = np.arange(np.floor(lo) + 1.0, np.ceil(hi))
                t.extend(((k - c0) / (c1 - c0)).tolist())
        t = np.unique(np.clip(np.asarray(t), 0.0, 1.0))
        px = x0 + t * (x1 - x0)
        py = y0 + t * (y1 - y0)
        xs.append(0.5 * (px[1:] + px[:-1]))
        ys.append(0.5 * (py[1:] + py[:-1]))
        hs.append(py[1:] - py[:-1])
    if not xs:
        return np.zeros(0), np.zeros(0), np.zeros(0)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(hs)


def coverage_array(vertices, nx, ny, dx, origin=(0.0, 0.0)):
    """numpy (nx, ny) coverage of the polygon with world-unit `vertices`
    (cell (i, j) spans [i dx, (i+1) dx] x [j dx, (j+1) dx] from `origin`)."""
    v = (np.asarray(vertices, dtype=np.float64) - np.asarray(origin, dtype=np.float64)) / dx
    signed = 0.5 * float(np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1]))
    xm, ym, h = _pieces(v)
    acc = np.zeros((nx + 1, ny), dtype=np.float64)
    j = np.floor(ym).astype(np.int64)
    i = np.floor(xm).astype(np.int64)
    keep = (j >= 0) & (j < ny) & (i < nx)
    i, j, xm, h = i[keep], j[keep], xm[keep], h[keep]
    left = i < 0                              # whole height to every grid cell of the row
    np.add.at(acc, (np.zeros(int(left.sum()), dtype=np.int64), j[left]), h[left])
    i, j, xm, h = i[~left], j[~left], xm[~left], h[~left]
    np.add.at(acc, (i, j), h * ((i + 1) - xm))
    np.add.at(acc, (i + 1, j), h * (xm - i))
    cov = np.cumsum(acc[:nx], axis=0)
    # counter-clockwise edges go up on the right: the winding sum is -area there
    cov = -cov if signed > 0 else cov
    return snap(np.clip(cov, 0.0, 1.0))


def snap(cov):
    """The reference's snapping of nearly full / empty cells
    (raster.py:60-65): numpy isclose with atol 1e-9 and its default rtol
    1e-5, so cells within 1e-9 + 1e-5 of full become exactly 1 and cells
    within 1e-9 of empty exactly 0 (the eps maps must match the
    reference's, quirk included)."""
    cov[np.isclose(cov, 1.0, atol=1e-9)] = 1.0
    cov[np.isclose(cov, 0.0, atol=1e-9)] = 0.0
    return cov


def coverage(poly, grid, origin=(0.0, 0.0)):
    """Torch (nx, ny) float64 coverage of a Polygon on `grid`, on grid.device."""
    cov = coverage_array(poly.vertices, grid.nx, grid.ny, grid.dx, origin)
    return torch.tensor(cov, dtype=torch.float64, device=getattr(grid, "device", "cpu"))


def normals(cov):
    """Unit interface normals (pointing out of the shape: minus the coverage
    gradient, central differences with edge replication) on boundary cells
    (0 < cov < 1), zero elsewhere: (nx, ny, 2)."""
    p = torch.nn.functional.pad(cov[None, None], (1, 1, 1, 1), mode="replicate")[0, 0]
    g = torch.stack([-(p[2:, 1:-1] - p[:-2, 1:-1]) * 0.5, -(p[1:-1, 2:] - p[1:-1, :-2]) * 0.5], dim=-1)
    length = torch.linalg.norm(g, dim=-1, keepdim=True)
    edge = ((cov > 0) & (cov < 1)).unsqueeze(-1) & (length > 0)
    return torch.where(edge, g / length.clamp(min=1e-300), torch.zeros_like(g))
