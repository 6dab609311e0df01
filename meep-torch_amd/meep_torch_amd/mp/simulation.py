"""meep-compatible Simulation facade over the HIP FDTD engine (2D TM/TE).

Covers the hook-free flow of meep_torch/mp/simulation.py: construction
(:214-381), structure rasterization (:511-573), sources (:796-880),
add_flux/get_fluxes/get_flux_freqs (:1357-1386,1809-1870),
add_dft_fields/get_dft_array (:1592-1641), run(until | until_after_sources |
n_steps, step functions, stop conditions) (:1096-1142,1239-1302),
get_field_point (:983-1024), get_flux_data / load_flux_data /
load_minus_flux_data (:1882-1926) and get_array (:2119-2150).

3D cells (dimensions=3) take the reference's 3D adapter path, in sim3d.py.

Time stepping semantics (2D) follow the reference: every run() starts from zero
fields and zero DFT accumulators and integrates to the accumulated clock
(so consecutive runs reproduce one continuing run), in chunks
[this run's steps, previous clock] exactly as its _run_until loop issues
them (meep_torch/mp/simulation.py:1271-1295); each chunk's monitor sums are
added in chunk order.  Step functions and stop conditions (run_control.py)
split the run into engine calls at the points where a hook must look at the
fields.  Symmetries, k-points and geometry without a 2D cross-section
raise NotImplementedError.
"""
from collections import namedtuple

import numpy as np
import torch

from ..core import engine2d
from ..core.coeffs import build_te_coeffs, build_tm_coeffs
from ..core.grid import Grid
from ..core.monitors import DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY, FluxMonitorTM
from ..core.sources import LineSource
from ..core.stepper_te import fields_of as te_fields_of, init_te_state
from ..core.stepper_tm import fields_of as tm_fields_of, init_tm_state
from . import geometry as G
from . import run_control as RC
from .sim3d import Sim3D
from .sources import AUTOMATIC, EigenModeSource, port_line, source_span
from .symmetry import Folding, check_symmetry
from .vec import Vector3, pol_for

_DEFAULTS = dict(eps_averaging=True, symmetries=None, force_complex_fields=False, m=0, k_point=False,
                 bfast_scaled_k=None, material_function=None, epsilon_func=None, epsilon_input_file="",
                 output_volume=None, force_all_components=False, geometry_center=None, extra_materials=None,
                 kz_2d="complex", progress_interval=4, subpixel_tol=1e-4, subpixel_maxeval=100000,
                 loop_tile_base_db=0, loop_tile_base_eh=0, ensure_periodicity=True, num_chunks=0,
                 accurate_fields_near_cylorigin=False, filename_prefix=None, output_single_precision=False,
                 split_chunks_evenly=True, chunk_layout=None, collect_stats=False)


class FluxRegion:
    def __init__(self, center=Vector3(), size=Vector3(), direction=None, weight=1.0, volume=None):
        if volume is not None:
            center, size = volume.center, volume.size
        self.center, self.size = Vector3(*center), Vector3(*size)
        self.direction = direction
        self.weight = weight


class Volume:
    def __init__(self, center=Vector3(), size=Vector3()):
        self.center, self.size = Vector3(*center), Vector3(*size)


class _Dft:
    def __init__(self, sim, name, freqs):
        self.sim, self.name, self.freqs = sim, name, list(freqs)
        self.nfreqs = len(self.freqs)


class DftFlux(_Dft):
    pass


class DftFields(_Dft):
    pass


def _name(f):
    return f.name if isinstance(f, _Dft) else f


# meep's get_eigenmode_coefficients result (meep_torch/mp/run_control.py:193-195)
EigCoeffsResult = namedtuple("EigCoeffsResult", ["alpha", "vgrp", "kpoints", "kdom", "cscale"])


class _FieldProbe:
    """Engine point probe of one state component (reduce 'stack')."""
    reduce = "stack"

    def __init__(self, name, ix, iy):
        self.name, self.ix, self.iy = name, ix, iy

    def engine_terms(self, f):
        from ..core.monitors import _ids
        fid = _ids(f).get(self.name)
        sem = (fid, 4, self.ix, self.iy) if fid is not None else None
        return "probe", [(f[self.name][self.ix, self.iy].reshape(1, 1), None, 0, sem)]


class _SourceSum:
    """Several facade sources driving one run (the engine sums
    amp_i(t) * mask_i every step, engine2d's multi-source plan)."""

    def __init__(self, parts):
        self.parts = list(parts)


class Simulation(Sim3D):
    def __init__(self, *, cell_size, resolution, geometry=None, sources=None, boundary_layers=None,
                 default_material=None, Courant=0.5, dtype=torch.float64, device=None, dimensions=2,
                 **options):
        for k, v in options.items():
            if k not in _DEFAULTS:
                raise TypeError(f"Simulation() got an unexpected keyword argument {k!r}")
            if k == "geometry_center" and (v is None or Vector3(*v) == Vector3()):
                continue
            if k == "symmetries":
                continue
            if k == "eps_averaging" or v == _DEFAULTS[k]:
                continue
            raise NotImplementedError(f"Simulation({k}=...) is not supported by the HIP FDTD facade")
        self._fold = Folding((), None, "tm", None, None)           # replaced when a 2D structure is built
        self._cut = {}
        if dimensions == 3:
            if options.get("symmetries"):
                raise NotImplementedError("Mirror / Rotate2 / Rotate4 symmetries are 2D-only")
            self._init3(cell_size, resolution, geometry, sources, boundary_layers, default_material, Courant, dtype,
                        device)
            return
        if dimensions != 2:
            raise NotImplementedError(f"the Simulation facade drives 2D and 3D cells, not dimensions={dimensions}")
        self.dimensions = 2
        self.symmetries = list(options.get("symmetries") or [])   # meep-mutable until the structure is built
        for sym in self.symmetries:
            check_symmetry(sym)
        self.dtype = dtype
        self.cell_size = Vector3(*cell_size)
        self.resolution = resolution
        self.gs = G.grid_spec(self.cell_size, resolution)
        kw = {} if device is None else {"device": device}
        self.grid = Grid(nx=self.gs.nx, ny=self.gs.ny, dx=self.gs.dx, dtype=dtype, courant=float(Courant), **kw)
        self._sources = list(sources or [])
        self._pol = pol_for(self._sources[0].component) if self._sources else "tm"
        self._geometry = list(geometry or [])
        self._default = default_material if default_material is not None else G.Medium(epsilon=1.0)
        if isinstance(self._default, (int, float)):
            self._default = G.Medium(epsilon=float(self._default))
        self.boundary_layers = list(boundary_layers or [])
        self._npml_sides = G.pml_sides(self.boundary_layers, self.gs)
        self._monitors = {}        # name -> (kind, monitor, meta)
        self._flux_parts = {}
        self._acc = {}
        self._t_step = 0
        self._state = None
        self._raster = None
        self._coeffs = None
        self.run_index = 0
        # S-parameter path (mp/simulation.py:353-355,1305-1338,1766-1789)
        self._eig_source = next((q for q in self._sources if isinstance(q, EigenModeSource)), None)
        self._mode_ports = []
        self._mode_freqs = None
        self._port_index = {}
        self._S = None

    # ---- structure -----------------------------------------------------
    def _init_structure(self):
        if self.dimensions == 3:
            return self._init_structure3()
        if self._coeffs is not None:
            return
        pol = pol_for(self._sources[0].component) if self._sources else "tm"
        self._pol = pol
        self._raster = G.rasterize(self._geometry, self.gs, self._default.resolve())
        g = self.grid
        npml = (tuple(self._npml_sides[0]), tuple(self._npml_sides[1]))
        media = G.rasterize_media(self._geometry, self.gs, self._default)
        nabs = G.pml_sides(self.boundary_layers, self.gs, absorber=True)
        if any(any(p) for p in nabs):   # Absorber: the ramp joins sigma_d and sigma_b (mp/simulation.py:539-550)
            ramp = G.absorber_ramp(nabs, g.nx, g.ny, g.dx, g.dtype)
            sd, sb, mu = media["cond"] if media["cond"] is not None else (
                np.zeros((g.nx, g.ny)), np.zeros((g.nx, g.ny)), np.ones((g.nx, g.ny)))
            media["cond"] = (sd + ramp, sb + ramp, mu)
        # symmetries (mp/simulation.py:383-494,551-570): every map is
        # rasterised on the full cell, then the run keeps the part at and
        # above the cut plane(s) on a smaller grid closed by the cut walls
        self._fold = Folding(self.symmetries, self.gs, pol, self._npml_sides, nabs)
        cut = {}
        if self._fold.active:
            cx, cy = self._fold.cx, self._fold.cy
            half = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a)[cx:, cy:])
            self._raster = tuple(half(a) for a in self._raster)
            media = dict(media, masks=[half(m) for m in media["masks"]], chi3=half(media["chi3"]),
                         cond=None if media["cond"] is None else tuple(half(a) for a in media["cond"]))
            self.grid = g = Grid(nx=self.gs.nx - cx, ny=self.gs.ny - cy, dx=self.gs.dx, dtype=self.dtype,
                                 courant=g.courant, device=g.device)
            cut = self._fold.cut_kwargs(pol)
        self._cut = cut
        has_media = bool(media["poles"]) or media["chi3"] is not None or media["cond"] is not None
        self._n_poles = len(media["poles"])
        if pol == "tm":
            eps = G.to_tensor(self._raster[0], g.dtype, g.device)
            kw = {}
            if media["poles"]:
                kw.update(poles=media["poles"], pole_masks=[G.to_tensor(m, g.dtype, g.device) for m in media["masks"]])
            if media["chi3"] is not None:
                kw["chi3"] = G.to_tensor(media["chi3"], g.dtype, g.device)
            if media["cond"] is not None:
                sd, sb, mu = (G.to_tensor(v, g.dtype, g.device) for v in media["cond"])
                kw.update(sigma_d=sd, sigma_b=sb, mu=mu)
            self._coeffs = build_tm_coeffs(g, eps_inf=eps, npml=npml, **kw, **cut)
        else:
            kw = {}
            if media["chi3"] is not None:
                kw["chi3"] = G.to_tensor(media["chi3"], torch.float64, g.device)
            if media["poles"]:
                kw.update(poles=media["poles"],
                          pole_masks=[G.to_tensor(m, torch.float64, g.device) for m in media["masks"]])
            if media["cond"] is not None:
                sd, sb, mu = (G.to_tensor(v, torch.float64, g.device) for v in media["cond"])
                kw.update(sigma_d=sd, sigma_b=sb, mu=mu)
            eps_xx = G.to_tensor(self._raster[1], torch.float64, g.device)
            eps_yy = G.to_tensor(self._raster[2], torch.float64, g.device)
            c = build_te_coeffs(g, eps_inf=eps_xx, npml=npml, **kw, **cut)
            self._coeffs = c._replace(eps_y=(0.5 * (eps_yy[:, :-1] + eps_yy[:, 1:])).to(g.dtype))

    def init_sim(self):
        self._init_structure()

    def _make_sources(self):
        out = []
        for spec in self._sources:
            ix, j0, j1 = source_span(spec, self.gs)
            if self._fold.active:
                # a source below a cut is dropped, not folded: the cut wall
                # supplies its image (mp/simulation.py:822-837); a span that
                # crosses the plane keeps its part above it
                ix, mirx = self._fold.foldpt(ix, "x")
                j0, j1, _, miry = self._fold.fold(j0, j1, "y")
                if mirx or miry:
                    continue
            src = LineSource(self.grid, ix, j0, j1, profile=spec.src.profile())
            src.mask = src.mask * spec.amplitude
            if self._pol == "te":
                src.mask = src.mask[:, :-1].contiguous()
            out.append(src)
        return out

    def _source_cutoff(self):
        return max((float(getattr(s.src.profile(), "t_end", 0.0)) for s in self._sources), default=0.0)

    # ---- monitors ------------------------------------------------------
    def _line_monitor(self, region, freqs):
        c, s, gs = region.center, region.size, self.gs
        x_normal = abs(s.x) <= abs(s.y)
        if x_normal:
            idx = G.cell_index(c.x, gs)
            y0, y1 = c.y - s.y / 2 + gs.sy / 2, c.y + s.y / 2 + gs.sy / 2
            lo, hi = G.span_index_range(max(0.0, y0), min(gs.sy, y1), gs, "y")
        else:
            idx = max(0, min(gs.ny, G.cell_index_y(c.y, gs)))
            x0, x1 = c.x - s.x / 2 + gs.sx / 2, c.x + s.x / 2 + gs.sx / 2
            lo, hi = G.span_index_range(max(0.0, x0), min(gs.sx, x1), gs, "x")
        meta = {"idx": idx, "lo": lo, "hi": hi, "wt": complex(region.weight or 1.0).real}
        if self._fold.active:
            # fold the line onto the kept part (mp/simulation.py:1829-1843): a
            # span crossing the plane keeps its upper part and doubles back
            # (mir2); a line below a cut maps onto its image, where the
            # mirrored tangential H flips against E (msign)
            idx, pos_mir = self._fold.foldpt(idx, "x" if x_normal else "y")
            lo, hi, meta["mir2"], span_mir = self._fold.fold(lo, hi, "y" if x_normal else "x")
            self._fold.rot_guard("a flux/mode region", pos_mir, meta["mir2"], span_mir)
            if pos_mir:
                meta["msign"] = -1.0
            meta.update(idx=idx, lo=lo, hi=hi)
        if self._pol == "tm":
            mon = FluxMonitorTM(self.grid, freqs, x_normal=x_normal, idx=idx, lo=lo, hi=hi)
        elif x_normal:
            mon = FluxMonitorTE(self.grid, freqs, ix=idx, lo=lo, hi=hi)
        else:
            mon = FluxMonitorTEY(self.grid, freqs, iy=idx, lo=lo, hi=hi)
        return mon, meta

    def add_flux(self, fcen, df, nfreq, *regions, name=None):
        freqs = np.linspace(fcen - df / 2.0, fcen + df / 2.0, nfreq).tolist()
        self._init_structure()
        if self.dimensions == 3:
            return DftFlux(self, self._add_flux3(freqs, list(regions), name), freqs)
        name = name or f"flux{len(self._monitors)}"
        parts = []
        for r, region in enumerate(regions):
            mon, meta = self._line_monitor(region, freqs)
            key = name if len(regions) == 1 else f"{name}:r{r}"
            self._monitors[key] = ("flux", mon, meta)
            parts.append(key)
        if len(regions) > 1:
            self._flux_parts[name] = parts
        return DftFlux(self, name, freqs)

    def add_mode_monitor(self, fcen, df, nfreq, *regions, name=None):
        """meep add_mode_monitor (mp/simulation.py:1766-1789): with an
        EigenModeSource a mode port of the S-parameter run, otherwise a flux
        monitor."""
        if self.dimensions == 3:
            raise NotImplementedError("3D EigenModeSource / S-parameters are not supported")
        if len(regions) != 1:
            raise NotImplementedError("add_mode_monitor takes exactly one region")
        if self._eig_source is None:
            return self.add_flux(fcen, df, nfreq, *regions, name=name)
        freqs = np.linspace(fcen - df / 2.0, fcen + df / 2.0, nfreq).tolist()
        region = regions[0]
        name = name or f"mode{len(self._mode_ports)}"
        self._mode_ports.append((name,) + port_line(region.center, region.size, self.gs) + (region.direction,))
        self._mode_freqs = freqs
        return DftFlux(self, name, freqs)

    def _run_modes(self, n_steps):
        """mp/simulation.py:1304-1338: the EigenModeSource's port is port 0;
        each mode monitor becomes a port (one on the source's own line maps
        to port 0, the reflection), and the run is compute_smatrix's
        per-port launch sequence, kept as S."""
        from ..ports import Port, compute_smatrix
        if self._fold.active:
            raise NotImplementedError("EigenModeSource S-parameters with a symmetry are not supported")
        es = self._eig_source
        c = self._coeffs
        if self._n_poles or (c.d1 if self._pol == "tm" else c.d1x) is not None or c.has_kerr:
            raise NotImplementedError("EigenModeSource S-parameters in a cell with dispersive, nonlinear or "
                                      "conductive media are not supported")
        if es.eig_band != 1:
            raise NotImplementedError("EigenModeSource eig_band > 1: the S-parameter path launches band 1")
        axis, idx, lo, hi = port_line(es.center, es.size, self.gs)
        ports = [Port(axis=axis, index=idx, lo=lo, hi=hi, direction=1 if es.direction == AUTOMATIC else es.direction)]
        self._port_index = {}
        for name, ax, i, l, h, d in self._mode_ports:
            if (ax, i, l, h) == (axis, idx, lo, hi):
                self._port_index[name] = 0
                continue
            self._port_index[name] = len(ports)
            ports.append(Port(axis=ax, index=i, lo=l, hi=h, direction=d))
        src = es.src
        fcen, fwidth = float(src.frequency), 1.0 / float(getattr(src, "width", 0) or 10.0)
        # TM: the rasterised eps; TE: its xx map, which compute_smatrix
        # averages onto the Ey nodes itself (mp/simulation.py:1335)
        eps = G.to_tensor(self._raster[0 if self._pol == "tm" else 1], self.grid.dtype, self.grid.device)
        npml = max(max(p) for p in self._npml_sides)
        self._S = compute_smatrix(eps, self.grid, ports, pol=self._pol, fcen=fcen, fwidth=fwidth,
                                  freqs=self._mode_freqs or [fcen], n_steps=n_steps, npml=npml).s

    def get_eigenmode_coefficients(self, flux, bands, eig_parity=0, eig_vol=None, eig_resolution=0,
                                   eig_tolerance=1e-12, kpoint_func=None, direction=AUTOMATIC):
        """meep get_eigenmode_coefficients on a mode port of an EigenModeSource
        run (mp/simulation.py:1928-1985), TM or TE (x-normal ports) cells:
        alpha (1, nfreq, 2) whose forward
        slot holds the S-parameter column S[:, port, 0] (the source port's
        own line reports the reflection) and whose backward slot is zero.
        The plane-wave decomposition of Bloch-periodic flux planes is not
        built."""
        for nm, v, ok in (("eig_parity", eig_parity, 0), ("eig_vol", eig_vol, None),
                          ("kpoint_func", kpoint_func, None)):
            if v != ok:
                raise NotImplementedError(f"get_eigenmode_coefficients {nm} is not supported")
        if direction not in (AUTOMATIC, None):
            raise NotImplementedError("get_eigenmode_coefficients direction: the monitor normal only")
        if not isinstance(bands, (list, tuple, range)):
            raise TypeError("get_eigenmode_coefficients: bands must be a list")
        name = _name(flux)
        if self._S is None or name not in self._port_index:
            raise NotImplementedError(f"get_eigenmode_coefficients({name!r}): only mode ports of an "
                                      "EigenModeSource run are decomposed (no plane-wave branch)")
        if list(bands) != [1]:
            raise NotImplementedError(f"get_eigenmode_coefficients(bands={list(bands)}): mode ports carry band 1")
        col = self._S[:, self._port_index[name], 0].detach().cpu().numpy()
        alpha = np.zeros((1, col.shape[0], 2), dtype=np.complex128)
        alpha[0, :, 0] = col
        return EigCoeffsResult(alpha, np.zeros(col.shape[0]), [], [], np.zeros(col.shape[0]))

    def add_dft_fields(self, components, fcen, df, nfreq, where=None, center=None, size=None, name=None,
                       yee_grid=False, decimation_factor=0):
        freqs = np.linspace(fcen - df / 2.0, fcen + df / 2.0, nfreq).tolist()
        if where is not None:
            center, size = where.center, where.size
        self._init_structure()
        name = name or f"dft{len(self._monitors)}"
        if self.dimensions == 3:
            region = None if center is None and size is None else (center, size)
            self._add_dft3(freqs, list(components), name, region)
            return DftFields(self, name, freqs)
        box = None
        if center is not None or size is not None:
            x0, y0, x1, y1 = G.rect_corners(Vector3(*(center or Vector3())), Vector3(*(size or Vector3())),
                                            self.gs)
            i0, i1 = G.span_index_range(x0, x1, self.gs, "x")
            j0, j1 = G.span_index_range(y0, y1, self.gs, "y")
            box = (i0, i1, j0, j1)
        for comp in components:
            if self._fold.active and comp not in ("Ez", "Dz"):
                raise NotImplementedError(f"add_dft_fields({comp}) with a symmetry: staggered components do not "
                                          f"de-mirror onto the full cell")
            mon = DFTFieldsMonitor(self.grid, freqs, component=comp)
            self._monitors[f"{name}:{comp}"] = ("dft", mon, {"freqs": freqs, "box": box})
        return DftFields(self, name, freqs)

    # ---- running -------------------------------------------------------
    def run(self, *step_funcs, until=None, until_after_sources=None, n_steps=None):
        self.run_index += 1
        self._init_structure()
        if self.dimensions == 3:
            if step_funcs or callable(until) or callable(until_after_sources) or \
                    isinstance(until, (list, tuple)) or isinstance(until_after_sources, (list, tuple)):
                raise NotImplementedError("step functions / stop conditions: 2D runs only")
            if n_steps is None:
                T = until if until is not None else (
                    self._source_cutoff() + until_after_sources if until_after_sources is not None else 0)
                n_steps = max(1, int(round(float(T) / self._coeffs.dt)))
            self._run3(int(n_steps))
            return
        if self._eig_source is not None:
            if step_funcs or callable(until) or callable(until_after_sources) or \
                    isinstance(until, (list, tuple)) or isinstance(until_after_sources, (list, tuple)):
                raise NotImplementedError("step functions / stop conditions with an EigenModeSource")
            if n_steps is None:
                T = until if until is not None else self._source_cutoff() + float(until_after_sources or 0.0)
                n_steps = max(1, int(round(float(T) / self.grid.dt)))
            self._run_modes(int(n_steps))
            self._t_step += int(n_steps)
            return
        self._run_until(step_funcs, until, until_after_sources, n_steps)

    def _run_until(self, step_funcs, until, until_after_sources, n_steps):
        """meep_torch/mp/simulation.py:1239-1302: every run starts from zero
        fields (and the loaded flux data, if any), hooks fire between engine
        calls sized by the hook schedule, stop conditions are fed after every
        step, and the step counter continues from the previous run's clock
        while the fields' own counter restarts at 0 (so consecutive runs
        reproduce one continuing run)."""
        mode = self._pol
        dt = self.grid.dt
        cutoff = self._source_cutoff()
        begin, per_step, scheduled = RC.plan_hooks(step_funcs, cutoff)
        conds, gate, n_total = None, 0.0, n_steps
        stop = until if until is not None else until_after_sources
        if callable(stop) or isinstance(stop, (list, tuple)):
            conds = list(stop) if isinstance(stop, (list, tuple)) else [stop]
            if until is None:
                gate = cutoff            # until_after_sources: conditions count after the cutoff
        elif n_total is None:
            T = float(until) if until is not None else cutoff + float(until_after_sources or 0.0)
            n_total = max(1, int(round(T / dt)))
        if mode == "tm":
            st = init_tm_state(self.grid, n_poles=self._n_poles, **self._cut)
            names = tm_fields_of(st)
        else:
            st = init_te_state(self.grid, n_poles=self._n_poles)
            names = te_fields_of(st)
        work = {f: getattr(st, f).contiguous() for f in names}
        srcs = self._make_sources()
        # several sources: one injection of sum_i amp_i(t) mask_i per step,
        # and the monitors fold every step into one running sum, as the
        # reference's per-step loop does (mp/cpp/core.py:584-589,648-664,802-820)
        multi = len(srcs) > 1
        src = _SourceSum(srcs) if multi else (srcs[0] if srcs else None)
        target = work["Dz"].shape if mode == "tm" else work["Dy"].shape
        plan = engine2d.source_plan(src, target, self.grid.device, self.grid.dtype)
        keys = list(self._monitors)
        mons = [self._monitors[k][1] for k in keys]
        seeds = getattr(self, "_acc_seed", {})
        acc = [seeds[k].clone() if k in seeds else None for k in keys]
        # stop_when_fields_decayed only compares at the end of its dt window:
        # a device probe records its field every step and the engine runs a
        # whole window per call (same stop step as feeding it every step)
        probes = self._decay_probes(conds, work) if conds is not None else None
        raw = None
        if (multi or probes is not None) and mons:
            # one running sum per monitor, started from the loaded flux data:
            # the reference folds every step into the total on these paths
            # (per-step calls), not chunk partials
            raw = engine2d.run_chunk(mode, work, self._coeffs, 0, 0, None, mons, raw=True)
            for (_, _, buf), a in zip(raw, acc):
                if a is not None:
                    buf.copy_(a.reshape(buf.shape))
        self._state = st._replace(**work)
        self._acc = {k: a for k, a in zip(keys, acc) if a is not None}
        start = self._t_step
        stop_at = None if n_total is None else start + int(n_total)
        f_t = 0                          # the fields' own step counter
        for g in begin:
            RC.eval_step_func(self, g, "step")
        RC.fire_scheduled(self, scheduled, start * dt)
        while stop_at is None or self._t_step < stop_at:
            t = self._t_step * dt
            left = stop_at - self._t_step if stop_at is not None else 1 << 30
            k = RC.chunk_size(t, dt, left, per_step, scheduled, conds is not None and probes is None)
            if probes is not None:
                k = min([k] + [g.steps_to_check(self._t_step, dt) for g in conds])
            pm = [p for p, _ in probes] if probes is not None else []
            if raw is not None:
                raw = engine2d.run_chunk(mode, work, self._coeffs, f_t, k, plan, mons + pm, reuse=raw, raw=True)
                outs = engine2d.finish_outputs(raw, work)
                acc = [a.clone() for a in outs[:len(mons)]]
            else:
                outs = engine2d.run_chunk(mode, work, self._coeffs, f_t, k, plan, mons + pm)
                for i, o in enumerate(outs[:len(mons)]):
                    acc[i] = o if acc[i] is None else acc[i] + o
            step_before = self._t_step
            f_t += k
            self._t_step = f_t
            self._state = st._replace(**work)
            self._acc = dict(zip(keys, acc))
            t = self._t_step * dt
            RC.fire_per_step(self, per_step, t, dt)
            RC.fire_scheduled(self, scheduled, t)
            if probes is not None:
                vals = [(o.detach().double().cpu() * scale).tolist() for o, (_, scale) in zip(outs[len(mons):], probes)]
                hit = [False] * len(conds)
                for q in range(k):
                    t_q = (step_before + q + 1) * dt
                    hit = [bool(g.observe(v[q], t_q)) for g, v in zip(conds, vals)]
                if t >= gate - 1e-9 and any(hit):
                    break
            elif conds is not None:
                hit = [bool(g(self)) for g in conds]     # every condition sees every step
                if t >= gate - 1e-9 and any(hit):
                    break
        for func in step_funcs:
            RC.eval_step_func(self, func, "finish")

    def _decay_probes(self, conds, work):
        """(probe monitor, scale) per stop_when_fields_decayed condition, or
        None when some condition is another callable (those are fed every
        step).  The probe reads the component get_field_point reads (D as
        eps E, B as H, a component the state lacks as the primary field) at
        the same cell."""
        if not conds or not all(isinstance(g, RC.stop_when_fields_decayed) for g in conds):
            return None
        out = []
        for g in conds:
            ix, iy, sgn = self._probe_cell(Vector3(*g.pt))
            name, scale = (g.c if isinstance(g.c, str) else ""), sgn
            if name.startswith("D") and name != "Dielectric":
                name = "E" + name[1:]
                scale = sgn * float(np.asarray(self.get_epsilon())[ix, iy])
            elif name.startswith("B"):
                name = "H" + name[1:]
            if name not in work:
                name = "Hz" if self._pol == "te" else "Ez"
            f = work[name]
            out.append((_FieldProbe(name, min(ix, f.shape[0] - 1), min(iy, f.shape[1] - 1)), scale))
        return out

    def restart_fields(self):
        self._state = None
        self._t_step = 0

    def round_time(self):
        if self.dimensions == 3:
            self._init_structure()
            return self._t_step * self._coeffs.dt
        return self._t_step * self.grid.dt

    meep_time = round_time

    # ---- readout ---------------------------------------------------------
    def get_fluxes(self, f):
        name = _name(f)
        parts = self._flux_parts.get(name)
        if parts:
            return [sum(v) for v in zip(*(self.get_fluxes(p) for p in parts))]
        kind, mon, meta = self._monitors[name]
        acc = self._acc.get(name)
        if acc is None:
            return [0.0] * int(mon.freqs.numel())
        vals = mon.flux(acc, self.gs.dx).detach().cpu().numpy() * meta.get("wt", 1.0) * meta.get("msign", 1.0)
        if meta.get("mir2"):
            vals = 2.0 * vals    # the span crossed the mirror plane: the kept part holds half the flux
        return vals.tolist()

    def get_flux_freqs(self, f):
        name = _name(f)
        if any(p[0] == name for p in getattr(self, "_mode_ports", ())):   # a mode port of the S-parameter path
            return list(self._mode_freqs)
        parts = self._flux_parts.get(name)
        return self._monitors[parts[0] if parts else name][1].freqs.detach().cpu().numpy().tolist()

    def get_field_point(self, c, pt):
        """meep_torch/mp/simulation.py:983-1024 (no symmetries): the value of
        component c in the cell holding point pt; D components read
        eps * E there, B components read H; a component the state does not
        carry reads the run's primary field (Ez for TM, Hz for TE)."""
        if self.dimensions == 3:
            raise NotImplementedError("get_field_point: 2D runs only")
        if self._state is None:
            return 0.0
        ix, iy, sgn = self._probe_cell(Vector3(*pt))
        name, scale = (c if isinstance(c, str) else ""), 1.0
        if name.startswith("D") and name != "Dielectric":
            name = "E" + name[1:]
            scale = float(np.asarray(self.get_epsilon())[ix, iy])
        elif name.startswith("B"):
            name = "H" + name[1:]
        f = getattr(self._state, name, None)
        if f is None:
            f = self._state.Hz if self._pol == "te" else self._state.Ez
        return sgn * scale * float(f[min(ix, f.shape[0] - 1), min(iy, f.shape[1] - 1)])

    def _probe_cell(self, pt):
        """(ix, iy, sign) of the engine cell holding meep point pt: folded
        onto the kept part under a symmetry, with the Ez phase of the fold
        (mp/simulation.py:995-1001)."""
        gs = self.gs
        ix = max(0, min(gs.nx - 1, G.cell_index(pt.x, gs)))
        iy = max(0, min(gs.ny - 1, G.cell_index_y(pt.y, gs)))
        if not self._fold.active:
            return ix, iy, 1.0
        ix, mirx = self._fold.foldpt(ix, "x")
        iy, miry = self._fold.foldpt(iy, "y")
        px, py = self._fold.parity("Ez")
        return ix, iy, (px if mirx else 1.0) * (py if miry else 1.0)

    def get_flux_data(self, flux):
        """meep_torch/mp/simulation.py:1882-1898: the monitor's accumulated
        transforms as FluxData(E, H), complex (nfreq, points), regions of a
        multi-region monitor concatenated along the points."""
        if self.dimensions == 3:
            raise NotImplementedError("get_flux_data: 2D runs only")
        name = _name(flux)
        es, hs = [], []
        for p in self._flux_parts.get(name) or [name]:
            if p not in self._acc:
                raise RuntimeError(f"get_flux_data({name!r}): no accumulated fields -- run() first")
            a = self._acc[p].detach().cpu().numpy()
            es.append(a[:, 0] + 1j * a[:, 1])
            hs.append(a[:, 2] + 1j * a[:, 3])
        return RC.FluxData(E=np.concatenate(es, axis=-1), H=np.concatenate(hs, axis=-1))

    def load_flux_data(self, flux, fdata, _scale=1.0):
        """meep_torch/mp/simulation.py:1900-1921: seed the monitor's
        accumulators with fdata; every later run() accumulates on top."""
        if self.dimensions == 3:
            raise NotImplementedError("load_flux_data: 2D runs only")
        name = _name(flux)
        E, H = np.asarray(fdata.E), np.asarray(fdata.H)
        seeds, off = {}, 0
        for p in self._flux_parts.get(name) or [name]:
            meta = self._monitors[p][2]
            n = meta["hi"] - meta["lo"]
            e, h = E[:, off:off + n], H[:, off:off + n]
            off += n
            a = _scale * np.stack([e.real, e.imag, h.real, h.imag], axis=1)
            seeds[p] = torch.as_tensor(a, dtype=self.grid.dtype, device=self.grid.device)
        if off != E.shape[-1]:
            raise ValueError(f"load_flux_data({name!r}): data has {E.shape[-1]} points, monitor has {off}")
        self._acc_seed = {**getattr(self, "_acc_seed", {}), **seeds}

    def load_minus_flux_data(self, flux, fdata):
        """meep_torch/mp/simulation.py:1923-1926: load_flux_data with the
        transforms negated, so the next run subtracts them."""
        self.load_flux_data(flux, fdata, _scale=-1.0)

    def get_dft_array(self, f, component, num_freq):
        key = f"{_name(f)}:{component}"
        kind, mon, meta = self._monitors[key]
        a = self._acc[key][num_freq].detach().cpu().numpy()
        out = a[0] + 1j * a[1]
        if self._fold.active:
            out = self._fold.demirror(out, self._fold.parity(component))
        if meta.get("box") is not None:
            i0, i1, j0, j1 = meta["box"]
            out = out[i0:i1, j0:j1]
        return out

    def get_array(self, component="Dielectric", center=None, size=None):
        if center is not None or size is not None:
            raise NotImplementedError("get_array sub-volumes are not supported")
        if self.dimensions == 3:
            return self._get_array3(component)
        if component == "Dielectric":
            self._init_structure()
            a = (self._raster[0] if self._pol == "tm" else self._raster[1]).copy()
            return self._fold.demirror(a) if self._fold.active else a
        if self._state is None:
            raise RuntimeError("run() before get_array(fields)")
        fold = self._fold if self._fold.active else None
        if component in ("Ez", "Ex") and self._pol == "tm":
            a = self._state.Ez.detach().cpu().numpy()
            return fold.demirror(a, fold.parity("Ez")) if fold else a
        if component in ("Hz", "Hy") and self._pol == "te":
            a = self._state.Hz.detach().cpu().numpy()
            return fold.demirror_te_hz(a) if fold else a
        f = getattr(self._state, component, None)
        if f is None:
            raise ValueError(f"unsupported component {component!r}")
        if fold:
            if self._pol != "tm" or component not in ("Dz", "Hx", "Hy"):
                raise NotImplementedError(f"get_array({component!r}) with a symmetry: the staggered component "
                                          f"does not rebuild onto the full cell")
            return fold.demirror(f.detach().cpu().numpy(), fold.parity(component))
        return f.detach().cpu().numpy()

    def get_epsilon(self):
        return self.get_array("Dielectric")


def get_fluxes(f):
    return f.sim.get_fluxes(f)


def get_flux_freqs(f):
    return f.sim.get_flux_freqs(f)
