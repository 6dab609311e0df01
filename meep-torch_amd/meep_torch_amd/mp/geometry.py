"""Grid spec, materials, geometric objects (Block, Ellipsoid, Cylinder,
Sphere, Prism) and their rasterization onto the Yee grid (the host-side setup
of the Simulation facade).  Axis-aligned Blocks take the exact rectangle
coverage; every other object is the polygon of its 2D cross-section on the
exact polygon coverage of geometry.raster.

Coordinates follow meep: the cell is centred on the origin; grid cell (i, j)
covers the corner-origin square [i dx, (i+1) dx) x [j dx, (j+1) dx)
(meep_torch/mp/coords.py:1-125).  A Block's coverage of a cell is its exact
area fraction (the reference computes the same fraction by clipping the
polygon strip by strip, meep_torch/geometry/raster.py:37-67, and snaps values
within 1e-9 of 0 or 1; both are applied here).  Objects are painted in list
order: TM eps = cov * eps_obj + (1 - cov) * eps_before
(meep_torch/geometry/smoothing.py:11-12, meep_torch/geometry/scene.py:53-101);
TE uses the diagonal-projected Kottke pair at interface cells
(meep_torch/geometry/smoothing.py:15-23; Kottke, Farjadpour & Johnson,
Phys. Rev. E 77, 036611 (2008)).
"""
import math
from dataclasses import dataclass

import numpy as np
import torch

from .vec import ALL, Vector3


@dataclass(frozen=True)
class GridSpec:
    nx: int
    ny: int
    dx: float
    sx: float
    sy: float


def grid_spec(cell, resolution):
    """cell_size + resolution -> GridSpec: dx = 1/R, n = max(3, round(s R))
    (meep_torch/mp/coords.py:34-50, 2D)."""
    c = Vector3(*cell)
    dx = 1.0 / float(resolution)
    nx = max(3, int(round(c.x / dx))) if c.x > 0 else 3
    ny = max(3, int(round(c.y / dx))) if c.y > 0 else 3
    return GridSpec(nx, ny, dx, c.x if c.x > 0 else nx * dx, c.y if c.y > 0 else ny * dx)


def _clip(v, lo, hi):
    return max(lo, min(hi, v))


def cell_index(x, gs):
    """World x -> cell index round((x + sx/2) / dx) (meep_torch/mp/coords.py:79-85)."""
    return int(round((x + gs.sx / 2.0) / gs.dx))


def cell_index_y(y, gs):
    return int(round((y + gs.sy / 2.0) / gs.dx))


def span_index_range(a, b, gs, axis):
    """Corner-origin span [a, b] -> half-open index range, at least one cell
    (meep_torch/mp/coords.py:112-121)."""
    n = gs.nx if axis == "x" else gs.ny
    i0 = max(0, int(round(a / gs.dx)))
    i1 = min(n, int(round(b / gs.dx)))
    if i1 <= i0:
        i1 = min(n, i0 + 1)
    return i0, i1


def rect_corners(center, size, gs):
    """Centre + size (world) -> corner-origin (x0, y0, x1, y1), infinite extents
    and overhangs clipped to the cell (meep_torch/mp/coords.py:63-76)."""
    if math.isinf(size.x):
        x0, x1 = 0.0, gs.sx
    else:
        x0 = _clip(center.x - size.x / 2.0 + gs.sx / 2.0, 0.0, gs.sx)
        x1 = _clip(center.x + size.x / 2.0 + gs.sx / 2.0, 0.0, gs.sx)
    if math.isinf(size.y):
        y0, y1 = 0.0, gs.sy
    else:
        y0 = _clip(center.y - size.y / 2.0 + gs.sy / 2.0, 0.0, gs.sy)
        y1 = _clip(center.y + size.y / 2.0 + gs.sy / 2.0, 0.0, gs.sy)
    return x0, y0, x1, y1


def pml_cells(thickness, gs):
    return max(0, int(round(thickness / gs.dx)))


def _iso(sigma, sigma_diag, sigma_offdiag, what):
    """One scalar sigma from meep's sigma / sigma_diag / sigma_offdiag
    (meep_torch/mp/material.py:63-83): anisotropic values raise."""
    if sigma_offdiag is not None and any(float(v) for v in Vector3(*sigma_offdiag)):
        raise NotImplementedError(f"{what}: anisotropic sigma_offdiag is not supported")
    if sigma:
        return float(sigma)
    d = Vector3(*(sigma_diag if sigma_diag is not None else Vector3()))
    if not (d.x == d.y == d.z):
        raise NotImplementedError(f"{what}: anisotropic sigma_diag is not supported")
    return float(d.x)


class LorentzianSusceptibility:
    """chi(f) = sigma fn^2 / (fn^2 - f^2 - i f gamma); engine pole
    (fn, gamma, sigma) (meep_torch/mp/material.py:89-107)."""

    def __init__(self, frequency=0.0, gamma=0.0, sigma=None, sigma_diag=None, sigma_offdiag=None):
        self.frequency, self.gamma = float(frequency), float(gamma)
        self.sigma = _iso(sigma, sigma_diag, sigma_offdiag, "LorentzianSusceptibility")

    def to_pole(self):
        return (self.frequency, self.gamma, self.sigma)


class DrudeSusceptibility:
    """chi(f) = -sigma fp^2 / (f^2 + i f gamma); engine pole
    (0, gamma, sigma, fp): no restoring force, coupling at the plasma
    frequency (meep_torch/mp/material.py:110-132)."""

    def __init__(self, frequency=0.0, gamma=0.0, sigma=None, sigma_diag=None, sigma_offdiag=None):
        self.frequency, self.gamma = float(frequency), float(gamma)
        self.sigma = _iso(sigma, sigma_diag, sigma_offdiag, "DrudeSusceptibility")

    def to_pole(self):
        return (0.0, self.gamma, self.sigma, self.frequency)


class Medium:
    """Isotropic medium: epsilon (or index, eps = n^2), electric Lorentzian /
    Drude susceptibilities, chi3, D and B conductivities and mu (scalars).
    The media terms reach 2D TM runs; other meep Medium properties raise
    NotImplementedError by name."""

    def __init__(self, epsilon=None, index=None, E_susceptibilities=(), chi3=None, D_conductivity=None,
                 B_conductivity=None, mu=None, **unsupported):
        for k, v in unsupported.items():
            if v not in (None, 0, 0.0, (), [], 1, 1.0, False):
                raise NotImplementedError(f"Medium({k}=...) is not supported by the HIP FDTD path")
        if index is not None and epsilon is not None:
            raise ValueError("Medium: give epsilon or index, not both")
        self.epsilon = float(index) ** 2 if index is not None else (1.0 if epsilon is None else float(epsilon))
        for sus in E_susceptibilities or ():
            if not isinstance(sus, (LorentzianSusceptibility, DrudeSusceptibility)):
                raise NotImplementedError(f"susceptibility {type(sus).__name__} is not supported")
        self.E_susceptibilities = list(E_susceptibilities or ())
        self.chi3 = float(chi3 or 0.0)
        self.D_conductivity = float(D_conductivity or 0.0)
        self.B_conductivity = float(B_conductivity or 0.0)
        self.mu = float(mu or 1.0)

    def resolve(self):
        return self.epsilon

    def poles(self):
        return tuple(s.to_pole() for s in self.E_susceptibilities)

    def has_media(self):
        return bool(self.E_susceptibilities) or self.chi3 != 0.0 or self.D_conductivity != 0.0 or \
            self.B_conductivity != 0.0 or self.mu != 1.0


class GeometricObject:
    def __init__(self, material=None, center=Vector3(), epsilon_func=None):
        if epsilon_func is not None:
            raise NotImplementedError("epsilon_func geometry is not supported")
        self.material = material if material is not None else Medium()
        if isinstance(self.material, (int, float)):
            self.material = Medium(epsilon=float(self.material))
        self.center = Vector3(*center)

    def shift(self, v):
        raise NotImplementedError


def to_corner(v, gs):
    """World (centre-origin) point -> corner-origin (x, y), clipped to the
    cell, mp.inf mapped to the far edge (meep_torch/mp/coords.py:55-61)."""
    x = gs.sx if math.isinf(v.x) else v.x + gs.sx / 2.0
    y = gs.sy if math.isinf(v.y) else v.y + gs.sy / 2.0
    return _clip(x, 0.0, gs.sx), _clip(y, 0.0, gs.sy)


def circle_vertices(radius, cx, cy, gs, n=None):
    """The polygon that stands for a circle: n vertices from angle 0, n =
    round(2 pi r / dx) clamped to [16, 256] unless given (the reference's
    resolution rule, meep_torch/mp/geometry.py:47-51, kept so that the eps
    maps agree)."""
    if n is None:
        n = min(256, max(16, int(round(2 * math.pi * radius / gs.dx))))
    ang = np.linspace(0.0, 2 * math.pi, n, endpoint=False)
    return np.stack([cx + radius * np.cos(ang), cy + radius * np.sin(ang)], axis=1)


def polygon_coverage(vertices, gs):
    from ..geometry.polygons import Polygon
    from ..geometry.raster import coverage_array
    return coverage_array(Polygon(np.asarray(vertices)).vertices, gs.nx, gs.ny, gs.dx)


def _inplane_pair(e1, e2, e3):
    e1, e2, e3 = Vector3(*e1), Vector3(*e2), Vector3(*e3)
    if e1.z != 0.0 or e2.z != 0.0 or e3.x != 0.0 or e3.y != 0.0:
        raise NotImplementedError("Block basis must keep e1/e2 in the x-y plane (2D raster)")
    return e1.unit(), e2.unit()


class Block(GeometricObject):
    """Brick with basis e1 / e2 (in-plane rotation allowed): axis-aligned
    blocks take the exact rectangle coverage, rotated ones the polygon path."""

    def __init__(self, size, e1=Vector3(1, 0, 0), e2=Vector3(0, 1, 0), e3=Vector3(0, 0, 1), **kw):
        super().__init__(**kw)
        self.size = Vector3(*size)
        self.e1, self.e2, self.e3 = Vector3(*e1), Vector3(*e2), Vector3(*e3)

    def coverage(self, gs):
        u1, u2 = _inplane_pair(self.e1, self.e2, self.e3)
        if u1 == Vector3(1.0, 0.0, 0.0) and u2 == Vector3(0.0, 1.0, 0.0):
            x0, y0, x1, y1 = rect_corners(self.center, self.size, gs)
            return rect_coverage(x0, y0, x1, y1, gs)
        if math.isinf(self.size.x) or math.isinf(self.size.y):
            raise NotImplementedError("rotated Block with mp.inf size not supported")
        x, y = to_corner(self.center, gs)
        hx, hy = self.size.x / 2.0, self.size.y / 2.0
        return polygon_coverage([[x + a * u1.x + b * u2.x, y + a * u1.y + b * u2.y]
                                 for a, b in ((-hx, -hy), (hx, -hy), (hx, hy), (-hx, hy))], gs)


class Ellipsoid(Block):
    """Ellipse inscribed in its (possibly rotated) block, as a polygon of
    round(2 pi max(a, b) / dx) vertices clamped to [16, 256]."""

    def coverage(self, gs):
        u1, u2 = _inplane_pair(self.e1, self.e2, self.e3)
        x, y = to_corner(self.center, gs)
        a, b = self.size.x / 2.0, self.size.y / 2.0
        n = min(256, max(16, int(round(2 * math.pi * max(a, b) / gs.dx))))
        ang = np.linspace(0.0, 2 * math.pi, n, endpoint=False)
        ca, sb = a * np.cos(ang), b * np.sin(ang)
        return polygon_coverage(np.stack([x + ca * u1.x + sb * u2.x, y + ca * u1.y + sb * u2.y], axis=1), gs)


class Cylinder(GeometricObject):
    """Cylinder along z: its circular cross-section (n_pts vertices when
    given, else the circle rule of circle_vertices)."""

    def __init__(self, radius, axis=Vector3(0, 0, 1), height=1e20, n_pts=None, **kw):
        super().__init__(**kw)
        self.radius, self.axis, self.height, self.n_pts = float(radius), Vector3(*axis), float(height), n_pts

    def coverage(self, gs):
        if self.axis.x != 0.0 or self.axis.y != 0.0:
            raise NotImplementedError("Cylinder axis off z not supported (2D raster)")
        x, y = to_corner(self.center, gs)
        n = max(3, int(self.n_pts)) if self.n_pts is not None else None
        return polygon_coverage(circle_vertices(self.radius, x, y, gs, n), gs)


class Sphere(GeometricObject):
    """Sphere: its x-y cross-section through the centre."""

    def __init__(self, radius, **kw):
        super().__init__(**kw)
        self.radius = float(radius)

    def coverage(self, gs):
        x, y = to_corner(self.center, gs)
        return polygon_coverage(circle_vertices(self.radius, x, y, gs), gs)


class Prism(GeometricObject):
    """Prism along z from world-coordinate vertices; a given centre moves the
    vertices' centroid onto it (meep's rule).  Vertices map to the cell as
    points do (clipped to it)."""

    def __init__(self, vertices, height=1e20, axis=Vector3(0, 0, 1), center=None, sidewall_angle=0, **kw):
        if sidewall_angle != 0:
            raise NotImplementedError("Prism sidewall_angle not supported (2D raster)")
        ax = Vector3(*axis)
        if ax.x != 0.0 or ax.y != 0.0:
            raise NotImplementedError("Prism axis must be along z (2D raster)")
        vs = [Vector3(*v) for v in vertices]
        centroid = Vector3(sum(v.x for v in vs) / len(vs), sum(v.y for v in vs) / len(vs),
                           sum(v.z for v in vs) / len(vs))
        if center is not None:
            d = Vector3(*center) - centroid
            vs = [v + d for v in vs]
            centroid = Vector3(*center)
        super().__init__(center=centroid, **kw)
        self.vertices, self.height, self.axis = vs, float(height), ax

    def coverage(self, gs):
        return polygon_coverage([to_corner(v, gs) for v in self.vertices], gs)


def _overlap_1d(a, b, n, dx):
    lo = np.arange(n, dtype=np.float64) * dx
    hi = (np.arange(n, dtype=np.float64) + 1.0) * dx
    return np.clip(np.minimum(hi, b) - np.maximum(lo, a), 0.0, None) / dx


def rect_coverage(x0, y0, x1, y1, gs):
    """Area fraction of every cell covered by [x0, x1] x [y0, y1] (corner
    coordinates), clipped to [0, 1] and snapped as the reference's polygon
    coverage snaps (geometry.raster.snap)."""
    from ..geometry.raster import snap
    cov = np.outer(_overlap_1d(x0, x1, gs.nx, gs.dx), _overlap_1d(y0, y1, gs.ny, gs.dx))
    return snap(np.clip(cov, 0.0, 1.0))


def interface_normals(cov):
    """Unit normals (pointing out of the object) from central differences of
    the coverage with edge replication, on interface cells (0 < cov < 1)
    only; zero elsewhere."""
    p = np.pad(cov, 1, mode="edge")
    gx = 0.5 * (p[2:, 1:-1] - p[:-2, 1:-1])
    gy = 0.5 * (p[1:-1, 2:] - p[1:-1, :-2])
    length = np.sqrt(gx * gx + gy * gy)
    ok = (cov > 0) & (cov < 1) & (length > 0)
    safe = np.where(ok, length, 1.0)
    return np.where(ok, -gx / safe, 0.0), np.where(ok, -gy / safe, 0.0)


def rasterize(objects, gs, background=1.0):
    """-> (eps_tm, eps_te_xx, eps_te_yy) float64 numpy arrays (nx, ny)."""
    eps_tm = np.full((gs.nx, gs.ny), float(background))
    te_xx = eps_tm.copy()
    te_yy = eps_tm.copy()
    for obj in objects:
        if not hasattr(obj, "coverage"):
            raise NotImplementedError(f"geometry object {type(obj).__name__} is not supported")
        cov = obj.coverage(gs)
        val = obj.material.resolve()
        old = eps_tm
        eps_tm = cov * val + (1.0 - cov) * old
        arith = eps_tm
        harm = 1.0 / (cov / val + (1.0 - cov) / old)
        nxn, nyn = interface_normals(cov)
        n2x, n2y = nxn * nxn, nyn * nyn
        edge = (n2x + n2y) > 0.5
        te_xx = np.where(edge, n2x * harm + (1.0 - n2x) * arith, arith)
        te_yy = np.where(edge, n2y * harm + (1.0 - n2y) * arith, arith)
    return eps_tm, te_xx, te_yy


def rasterize_media(objects, gs, background):
    """Media maps of the painter's scene (meep_torch/geometry/scene.py:53-101,
    meep_torch/mp/geometry.py:265-286, meep_torch/mp/simulation.py:526-550),
    float64 numpy (nx, ny):

    * poles / pole masks: the background's poles cover everything; each
      object scales every existing mask by (1 - coverage) and adds its
      coverage to the masks of its own poles (a pole equal to an earlier one
      shares its mask);
    * chi3: coverage-weighted over the objects whose chi3 is nonzero, then
      the background's chi3 (when nonzero) painted over the whole cell;
    * sigma_d / sigma_b / mu: when any object or the background has one,
      each is the arithmetic coverage blend of every object's value over the
      background's (objects without the property paint its default).
    Returns {"poles", "masks", "chi3" (or None), "cond" ((sd, sb, mu) or None)}."""
    shape = (gs.nx, gs.ny)
    poles = list(background.poles())
    masks = [np.ones(shape) for _ in poles]
    chi3, any_chi3 = np.zeros(shape), False
    covs = []
    for obj in objects:
        if not hasattr(obj, "coverage"):
            raise NotImplementedError(f"geometry object {type(obj).__name__} is not supported")
        cov = obj.coverage(gs)
        m = obj.material
        covs.append((cov, m))
        for k in range(len(masks)):
            masks[k] = (1.0 - cov) * masks[k]
        for p in m.poles():
            if p in poles:
                k = poles.index(p)
                masks[k] = cov + masks[k]
            else:
                poles.append(p)
                masks.append(cov.copy())
        if m.chi3 != 0.0:
            chi3 = cov * m.chi3 + (1.0 - cov) * chi3
            any_chi3 = True
    if background.chi3 != 0.0:
        full = np.ones(shape)
        chi3 = full * background.chi3 + (1.0 - full) * chi3
        any_chi3 = True
    cond = None
    bg = (background.D_conductivity, background.B_conductivity, background.mu)
    if any((m.D_conductivity, m.B_conductivity, m.mu) != (0.0, 0.0, 1.0) for _, m in covs) or bg != (0.0, 0.0, 1.0):
        maps = []
        for q, name in enumerate(("D_conductivity", "B_conductivity", "mu")):
            v = np.full(shape, bg[q])
            for cov, m in covs:
                v = cov * getattr(m, name) + (1.0 - cov) * v
            maps.append(v)
        cond = tuple(maps)
    return {"poles": poles, "masks": masks, "chi3": chi3 if any_chi3 else None, "cond": cond}


class PML:
    """Perfectly matched layer of `thickness` on `direction` (ALL, X, Y) and
    `side` (ALL, Low, High)."""

    def __init__(self, thickness, direction=ALL, side=ALL, R_asymptotic=1e-15, mean_stretch=1.0,
                 pml_profile=None):
        if pml_profile is not None or mean_stretch != 1.0:
            raise NotImplementedError("PML profile/stretch options are not supported")
        self.thickness = float(thickness)
        self.direction = direction
        self.side = side


class Absorber(PML):
    """meep mp.Absorber: the same edge selection as PML, realised as a
    matched sigma_d = sigma_b quadratic conductivity ramp (R 1e-15) instead
    of a CPML (meep_torch/mp/material.py:502-505, mp/simulation.py:539-550,
    615-627)."""


def pml_sides(layers, gs, absorber=False):
    """[[x_lo, x_hi], [y_lo, y_hi]] cell counts, max over the PML layers (or,
    absorber=True, over the Absorber layers)."""
    sides = [[0, 0], [0, 0]]
    for bl in layers:
        if not isinstance(bl, PML):
            raise NotImplementedError(f"boundary layer {type(bl).__name__} is not supported")
        if isinstance(bl, Absorber) != bool(absorber):
            continue
        n = pml_cells(bl.thickness, gs)
        axes = (0, 1) if bl.direction == ALL else (bl.direction,)
        for ax in axes:
            if ax not in (0, 1):
                raise ValueError("2D cell: PML direction must be X, Y or ALL")
            for s in ((0, 1) if bl.side == ALL else (bl.side,)):
                sides[ax][s] = max(sides[ax][s], n)
    return sides


def absorber_ramp(nabs, nx, ny, dx, dtype):
    """The 2D Absorber conductivity (mp/simulation.py:615-627), float64 numpy
    (nx, ny): max of the x and y absorber_sigma ramps, each computed in the
    grid's dtype."""
    from ..core.pml import absorber_sigma
    rx = absorber_sigma(nx, tuple(nabs[0]), dx, dtype=dtype)[:, None].expand(nx, ny)
    ry = absorber_sigma(ny, tuple(nabs[1]), dx, dtype=dtype)[None, :].expand(nx, ny)
    return torch.maximum(rx, ry).to(torch.float64).numpy()


def to_tensor(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).to(device=device, dtype=dtype)
