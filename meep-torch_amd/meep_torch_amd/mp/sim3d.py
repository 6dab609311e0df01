"""3D branch of the meep-compatible Simulation facade (dimensions=3).

Follows the reference's 3D adapter, meep_torch/mp/simulation.py:630-766 and
its 3D dispatch points (:1116-1137, :1372-1375, :1609-1629, :1860, :2119-2127):

* structure: eps, D_conductivity and B_conductivity of the default material,
  axis-aligned Blocks voxel-filled over their index box (coords.box3d,
  meep_torch/mp/coords.py:97-111; no subpixel averaging in 3D),
  build_3d_coeffs with per-side CPML counts; Lorentz / Drude
  susceptibilities as ADE poles (the reference's 3D adapter raises on
  them); chi3 and mu raise;
* sources: the first source only (:678-714): a zero-size source is a
  PointSource3D, a plane source is a PlaneSource along the axis of its
  smallest extent, full-plane when a transverse extent is infinite,
  otherwise windowed to [npml, min(n) - npml); Source.amplitude scales the
  waveform (:181-190);
* add_flux: exactly one region, a FluxMonitor3D normal to the region's
  smallest extent, aperture npml + 1 for a full-cell region or an explicit
  window (and no aperture) for a finite one (:730-766);
* add_dft_fields: full-cell DFTFieldsMonitor3D per component; regions raise;
* run: Courant is fixed at 0.5 (dt = 0.5 dx / sqrt(3)); every run starts
  from fresh fields (init_3d_state) and advances the clock by its own step
  count (:716-726, :1134-1137), as the reference does.
"""
import math

import torch

from ..core.monitors_3d import DFTFieldsMonitor3D, FluxMonitor3D, PlaneSource, PointSource3D, run_3d
from ..core.stepper_3d import build_3d_coeffs, init_3d_state
from . import geometry as G
from .vec import ALL, Vector3


class GridSpec3:
    def __init__(self, nx, ny, nz, dx, sx, sy, sz):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.sx, self.sy, self.sz = sx, sy, sz


def grid_spec3(cell, resolution):
    """cell_size + resolution -> 3D grid: dx = 1/R, n = max(3, round(s R))
    on every axis (meep_torch/mp/coords.py:34-48)."""
    c = Vector3(*cell)
    dx = 1.0 / float(resolution)
    n = [max(3, int(round(s / dx))) if s > 0 else 3 for s in (c.x, c.y, c.z)]
    s = [v if v > 0 else m * dx for v, m in zip((c.x, c.y, c.z), n)]
    return GridSpec3(n[0], n[1], n[2], dx, *s)


def cell_index3(v, gs):
    """Centre-origin world point -> (i, j, k) cell indices, cell-edge aligned
    (meep_torch/mp/coords.py:79-94)."""
    return (int(round((v.x + gs.sx / 2.0) / gs.dx)), int(round((v.y + gs.sy / 2.0) / gs.dx)),
            int(round((v.z + gs.sz / 2.0) / gs.dx)))


def box3d(center, size, gs):
    """Centre + size -> index box (i0, i1, j0, j1, k0, k1); infinite extents
    cover the axis (meep_torch/mp/coords.py:97-111)."""
    out = []
    for c, s, L in ((center.x, size.x, gs.sx), (center.y, size.y, gs.sy), (center.z, size.z, gs.sz)):
        if math.isinf(s):
            out.extend((0, int(round(L / gs.dx))))
        else:
            a = min(L, max(0.0, c - s / 2.0 + L / 2.0))
            b = min(L, max(0.0, c + s / 2.0 + L / 2.0))
            out.extend((max(0, int(round(a / gs.dx))), max(1, int(round(b / gs.dx)))))
    return tuple(out)


def pml_sides3(layers, gs, cell, absorber=False):
    """[[lo, hi]] CPML cell counts per axis (absorber=True: of the Absorber
    layers); direction ALL covers the axes the cell spans
    (meep_torch/mp/simulation.py:89-116, :359-379)."""
    sides = [[0, 0], [0, 0], [0, 0]]
    extent = (cell.x, cell.y, cell.z)
    for bl in layers:
        if not isinstance(bl, G.PML):
            raise NotImplementedError(f"boundary layer {type(bl).__name__} is not supported")
        if isinstance(bl, G.Absorber) != bool(absorber):
            continue
        n = G.pml_cells(bl.thickness, gs)
        if bl.direction == ALL:
            axes = tuple(a for a in range(3) if extent[a]) or (0, 1, 2)
        elif 0 <= bl.direction < 3:
            axes = (bl.direction,)
        else:
            raise ValueError(f"boundary layer direction {bl.direction} has no extent in a 3D cell")
        for ax in axes:
            for s in ((0, 1) if bl.side == ALL else (bl.side,)):
                sides[ax][s] = max(sides[ax][s], n)
    return sides


class _Amp:
    """A 3D source whose waveform is scaled by Source.amplitude."""

    def __init__(self, src, amp):
        self._src = src
        self.src = src.src
        self._amp = amp

    def amplitude(self, t_idx):
        return self._amp * self._src.amplitude(t_idx)


class Sim3D:
    """State and methods of a 3D Simulation (mixed into mp.Simulation)."""

    def _init3(self, cell_size, resolution, geometry, sources, boundary_layers, default_material, Courant, dtype,
               device):
        if float(Courant) != 0.5:
            raise NotImplementedError("3D Courant is fixed at 0.5 (the 3D stepper builds its own dt)")
        self.dimensions = 3
        self.dtype = dtype
        self.device = device or "cuda"
        self.cell_size = Vector3(*cell_size)
        self.resolution = resolution
        self.gs = grid_spec3(self.cell_size, resolution)
        self.grid = None
        self._sources = list(sources or [])
        self._geometry = list(geometry or [])
        self._default = default_material if default_material is not None else G.Medium(epsilon=1.0)
        if isinstance(self._default, (int, float)):
            self._default = G.Medium(epsilon=float(self._default))
        self.boundary_layers = list(boundary_layers or [])
        self._npml_sides = pml_sides3(self.boundary_layers, self.gs, self.cell_size)
        self._nabs_sides = pml_sides3(self.boundary_layers, self.gs, self.cell_size, absorber=True)
        self._npml = max(max(p) for p in self._npml_sides)
        self._monitors = {}
        self._flux_parts = {}
        self._acc = {}
        self._t_step = 0
        self._state = None
        self._eps3d = None
        self._coeffs = None
        self.run_index = 0

    # ---- structure -----------------------------------------------------
    @staticmethod
    def _check_medium3(m):
        if m.chi3 != 0.0 or m.mu != 1.0:
            raise NotImplementedError("chi3 and mu in 3D cells are not supported (the 3D stepper has neither)")

    def _init_structure3(self):
        """Voxel-filled Blocks in painter order (mp/simulation.py:634-677):
        eps, D_conductivity and B_conductivity volumes, conductivity passed
        only when some cell has it.  Beyond the reference's 3D adapter, which
        raises on them, Lorentz / Drude susceptibilities become ADE poles
        with one mask per (medium, pole): the cells that medium paints."""
        if self._coeffs is not None:
            return
        gs = self.gs
        shape = (gs.nx, gs.ny, gs.nz)
        bgm = self._default
        self._check_medium3(bgm)
        eps = torch.full(shape, float(bgm.resolve()), dtype=self.dtype, device=self.device)
        sd = torch.full(shape, bgm.D_conductivity, dtype=self.dtype, device=self.device)
        sb = torch.full(shape, bgm.B_conductivity, dtype=self.dtype, device=self.device)
        owner = torch.full(shape, -1, dtype=torch.int32, device=self.device)
        for q, obj in enumerate(self._geometry):
            if not isinstance(obj, G.Block):
                raise NotImplementedError(f"3D geometry {type(obj).__name__} is not supported (Block only)")
            m = obj.material
            self._check_medium3(m)
            i0, i1, j0, j1, k0, k1 = box3d(obj.center, obj.size, gs)
            eps[i0:i1, j0:j1, k0:k1] = float(m.resolve())
            sd[i0:i1, j0:j1, k0:k1] = m.D_conductivity
            sb[i0:i1, j0:j1, k0:k1] = m.B_conductivity
            owner[i0:i1, j0:j1, k0:k1] = q
        has_cond = bool((sd.abs().sum() + sb.abs().sum()).item())
        poles, masks = [], []
        for q, m in [(-1, bgm)] + list(enumerate(o.material for o in self._geometry)):
            if m.E_susceptibilities:
                cells = (owner == q).to(self.dtype)
                if bool(cells.any()):
                    for pole in m.poles():
                        poles.append(pole)
                        masks.append(cells)
        self._eps3d = eps
        self._n_poles = len(poles)
        self._coeffs = build_3d_coeffs(gs.nx, gs.ny, gs.nz, gs.dx, eps=eps,
                                       npml=tuple(tuple(p) for p in self._npml_sides),
                                       n_abs=tuple(tuple(p) for p in self._nabs_sides),
                                       sigma_d=sd if has_cond else None, sigma_b=sb if has_cond else None,
                                       poles=poles, pole_masks=masks or None,
                                       dtype=self.dtype, device=self.device)

    def _make_source3(self):
        if not self._sources:
            return None
        spec = self._sources[0]
        prof = spec.src.profile()
        comp = spec.component
        c = spec.center or Vector3()
        s = spec.size
        gs, dt = self.gs, self._coeffs.dt
        amp = float(spec.amplitude or 1.0)
        if s is None or (s.x == 0 and s.y == 0 and s.z == 0):
            i, j, k = cell_index3(c, gs)
            src = PointSource3D(dt, comp, i, j, k, prof)
        else:
            raw = [s.x or 0.0, s.y or 0.0, s.z or 0.0]
            sizes = [abs(v) for v in raw]
            axis = sizes.index(min(sizes))
            idx = cell_index3(c, gs)[axis]
            full_plane = any(math.isinf(t) for a, t in enumerate(raw) if a != axis)
            if full_plane:
                src = PlaneSource(gs.nx, gs.ny, gs.nz, dt, comp, axis, idx, prof, device=self.device)
            else:
                ap = self._npml
                src = PlaneSource(gs.nx, gs.ny, gs.nz, dt, comp, axis, idx, prof, lo=ap,
                                  hi=min(gs.nx, gs.ny, gs.nz) - ap, device=self.device)
        return _Amp(src, amp)

    # ---- monitors ------------------------------------------------------
    def _add_flux3(self, freqs, regions, name):
        if len(regions) != 1:
            raise NotImplementedError("3D add_flux takes exactly one region")
        region = regions[0]
        name = name or f"flux{len(self._monitors)}"
        gs = self.gs
        c = region.center or Vector3()
        s = region.size or Vector3()
        sizes = [abs(s.x or 0.0), abs(s.y or 0.0), abs(s.z or 0.0)]
        axis = sizes.index(min(sizes))
        normal = ("x", "y", "z")[axis]
        idx = cell_index3(c, gs)[axis]
        ap = self._npml + 1
        trans = [s.x or 0.0, s.y or 0.0, s.z or 0.0]
        others = [a for a in range(3) if a != axis]
        win = None
        if all(math.isfinite(trans[a]) for a in others):
            i0, i1, j0, j1, k0, k1 = box3d(c, s, gs)
            win = {"z": (i0, i1, j0, j1), "x": (j0, j1, k0, k1), "y": (i0, i1, k0, k1)}[normal]
            ap = 0
        mon = FluxMonitor3D(self._coeffs.dt, freqs, normal=normal, idx=idx, ap=ap, win=win, dtype=self.dtype)
        self._monitors[name] = ("flux", mon, {"normal": normal, "idx": idx})
        return name

    def _add_dft3(self, freqs, components, name, region):
        if region is not None:
            raise NotImplementedError("add_dft_fields with a region: 2D runs only")
        for comp in components:
            mon = DFTFieldsMonitor3D(self._coeffs.dt, freqs, component=comp, dtype=self.dtype)
            self._monitors[f"{name}:{comp}"] = ("dft", mon, {"freqs": freqs, "box": None})

    # ---- running -------------------------------------------------------
    def _run3(self, n_steps):
        gs = self.gs
        st = init_3d_state(gs.nx, gs.ny, gs.nz, npml=self._npml, n_poles=self._n_poles, dtype=self.dtype,
                           device=self.device)
        keys = list(self._monitors)
        mons = [self._monitors[k][1] for k in keys]
        st, outs = run_3d(st, self._coeffs, n_steps, source=self._make_source3(), monitors=mons)
        self._t_step += n_steps
        self._state = st
        self._acc = {k: a for k, a in zip(keys, outs)}

    def _get_array3(self, component):
        if component == "Dielectric":
            self._init_structure3()
            return self._eps3d.detach().cpu().numpy()
        if self._state is None:
            raise RuntimeError("run() before get_array(fields)")
        f = getattr(self._state, component, None)
        if f is None:
            raise ValueError(f"unsupported component {component!r}")
        return f.detach().cpu().numpy()


__all__ = ["Sim3D", "grid_spec3", "box3d", "cell_index3", "pml_sides3"]
