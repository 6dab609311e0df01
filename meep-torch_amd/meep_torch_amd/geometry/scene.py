"""Painter's-order compositing of polygon scenes into the engine's material
maps (the interface of meep_torch/geometry/scene.py:53-101: rasterize_scene
-> RasterizedMaterials, same fields and rules).

Items are (Polygon, material) in paint order; a material is an eps value or
an object with ``eps_inf`` and ``poles`` (a dispersive material).  Each item
blends over what is already painted by its coverage c:
  eps_tm   = c eps + (1 - c) eps_old                          (arithmetic)
  eps_te   = the diagonal Kottke pair of this item over eps_old: with the
             harmonic mean h and arithmetic mean a of the two, eps_xx =
             nx^2 h + (1 - nx^2) a on interface cells (unit normal n), a
             elsewhere; eps_yy likewise with ny (Kottke, Farjadpour &
             Johnson, Phys. Rev. E 77, 036611 (2008))
  pole masks: every existing mask scales by (1 - c); the item's own poles
             add c (a pole equal to an earlier one shares its mask)
  chi3     = c chi3 + (1 - c) chi3_old over the chi3 items."""
from typing import NamedTuple

import torch

from .raster import coverage, normals


class RasterizedMaterials(NamedTuple):
    eps_tm: torch.Tensor
    eps_te_xx: torch.Tensor
    eps_te_yy: torch.Tensor
    poles: list
    pole_masks: list
    chi3: torch.Tensor


def _eps_and_poles(m):
    if hasattr(m, "eps_inf"):
        return float(m.eps_inf), tuple(getattr(m, "poles", ()))
    return float(m), ()


def kottke_pair(eps_in, eps_out, cov, nrm):
    """(eps_xx, eps_yy) of a material eps_in covering `cov` of each cell over
    eps_out, with interface normals nrm (nx, ny, 2)."""
    arith = cov * eps_in + (1.0 - cov) * eps_out
    harm = 1.0 / (cov / eps_in + (1.0 - cov) / eps_out)
    n2x, n2y = nrm[..., 0] ** 2, nrm[..., 1] ** 2
    on_edge = (n2x + n2y) > 0.5
    return (torch.where(on_edge, n2x * harm + (1 - n2x) * arith, arith),
            torch.where(on_edge, n2y * harm + (1 - n2y) * arith, arith))


def rasterize_scene(items, grid, background=1.0, chi3_items=()):
    dev = getattr(grid, "device", "cpu")
    full = lambda v: torch.full((grid.nx, grid.ny), float(v), dtype=torch.float64, device=dev)
    bg_eps, bg_poles = _eps_and_poles(background)
    eps_tm, te_xx, te_yy = full(bg_eps), full(bg_eps), full(bg_eps)
    poles = list(bg_poles)
    masks = [full(1.0) for _ in bg_poles]
    for poly, mat in items:
        cov = coverage(poly, grid)
        val, own = _eps_and_poles(mat)
        prior = eps_tm
        eps_tm = cov * val + (1.0 - cov) * prior
        te_xx, te_yy = kottke_pair(val, prior, cov, normals(cov))
        masks = [(1.0 - cov) * m for m in masks]
        for p in own:
            if p in poles:
                k = poles.index(p)
                masks[k] = cov + masks[k]
            else:
                poles.append(p)
                masks.append(cov.clone())
    chi3 = full(0.0)
    for poly, val in chi3_items:
        cov = coverage(poly, grid)
        chi3 = cov * val + (1.0 - cov) * chi3
    return RasterizedMaterials(eps_tm=eps_tm, eps_te_xx=te_xx, eps_te_yy=te_yy, poles=poles, pole_masks=masks,
                               chi3=chi3)
