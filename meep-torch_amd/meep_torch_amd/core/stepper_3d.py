"""3D Yee problem/state types and single step (meep_torch/core/stepper_3d.py).

Coeffs3D / State3D keep the reference fields and shapes (State3D psi arrays are
full-grid; the kernels touch them only inside the CPML shells).  The HIP
stepper runs the CPML path, the matched-sigma absorber path of a run without
CPML, and the mixed CPML + absorber mode, with material conductivity
(sigma_d, sigma_b) and Lorentz / Drude ADE poles on the two-pass kernels; PMC
walls raise NotImplementedError."""
import math
from collections.abc import Mapping
from typing import NamedTuple

import torch

from .pml import absorber_sigma, cpml_profiles, side_counts
from .._engine import shell_bounds


class Coeffs3D(NamedTuple):
    dt: float
    dx: float
    eps: torch.Tensor
    damp: Mapping
    cpml: dict | None
    pols: dict | None = None
    pmc: tuple = ()
    shells: dict | None = None


class State3D(NamedTuple):
    Ex: torch.Tensor
    Ey: torch.Tensor
    Ez: torch.Tensor
    Hx: torch.Tensor
    Hy: torch.Tensor
    Hz: torch.Tensor
    psi_Ez_y: torch.Tensor | None = None
    psi_Ey_z: torch.Tensor | None = None
    psi_Ex_z: torch.Tensor | None = None
    psi_Ez_x: torch.Tensor | None = None
    psi_Ey_x: torch.Tensor | None = None
    psi_Ex_y: torch.Tensor | None = None
    psi_Hz_y: torch.Tensor | None = None
    psi_Hy_z: torch.Tensor | None = None
    psi_Hx_z: torch.Tensor | None = None
    psi_Hz_x: torch.Tensor | None = None
    psi_Hy_x: torch.Tensor | None = None
    psi_Hx_y: torch.Tensor | None = None
    Px_cur: torch.Tensor | None = None
    Px_prev: torch.Tensor | None = None
    Py_cur: torch.Tensor | None = None
    Py_prev: torch.Tensor | None = None
    Pz_cur: torch.Tensor | None = None
    Pz_prev: torch.Tensor | None = None


PSI_NAMES = ("psi_Ez_y", "psi_Ey_z", "psi_Ex_z", "psi_Ez_x", "psi_Ey_x", "psi_Ex_y",
             "psi_Hz_y", "psi_Hy_z", "psi_Hx_z", "psi_Hz_x", "psi_Hy_x", "psi_Hx_y")
COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _avg(v, axis):
    n = v.shape[axis]
    return 0.5 * (v.narrow(axis, 0, n - 1) + v.narrow(axis, 1, n - 1))


def _bc(v, ax):
    sh = [1, 1, 1]
    sh[ax] = v.shape[0]
    return v.reshape(sh)


def _axis_sides3(v):
    if isinstance(v, (tuple, list)):
        return tuple(side_counts(p) for p in v)
    return (side_counts(v),) * 3


class _Damp(Mapping):
    """Per-component (d1, d2) conductivity pairs of stepper_3d.py:217-235,
    built on first access: d1 = (1 - s dt/2) / (1 + s dt/2),
    d2 = dt / (1 + s dt/2) of sigma_d averaged along the E component's axis
    and of sigma_b averaged along the two axes across an H component
    (all-ones / all-dt at sigma = 0).  ``active`` tells whether material
    conductivity was given."""

    def __init__(self, eps, dt, sigma_d=None, sigma_b=None, sig_bound=None):
        self._eps, self._dt, self._cache = eps, dt, {}
        self._sd, self._sb, self._bound = sigma_d, sigma_b, sig_bound
        self.active = sigma_d is not None or sigma_b is not None or sig_bound is not None

    def _pair(self, key):
        e = self._eps
        src = self._sd if key[0] == "e" else self._sb
        s = (src.to(dtype=e.dtype, device=e.device) if src is not None else torch.zeros_like(e))
        if self._bound is not None:   # boundary absorber sigma + material sigma (stepper_3d.py:224-225)
            s = self._bound.to(dtype=e.dtype, device=e.device) + s
        dt = self._dt
        avg = {"ex": lambda v: _avg(v, 0), "ey": lambda v: _avg(v, 1), "ez": lambda v: _avg(v, 2),
               "hx": lambda v: _avg(_avg(v, 1), 2), "hy": lambda v: _avg(_avg(v, 0), 2),
               "hz": lambda v: _avg(_avg(v, 0), 1)}[key](s)
        return ((1 - avg * dt / 2) / (1 + avg * dt / 2), dt / (1 + avg * dt / 2))

    def __getitem__(self, key):
        if key not in self._cache:
            self._cache[key] = self._pair(key)
        return self._cache[key]

    def __iter__(self):
        return iter(("ex", "ey", "ez", "hx", "hy", "hz"))

    def __len__(self):
        return 6


def build_3d_coeffs(nx: int, ny: int, nz: int, dx: float, eps=None, npml=0, n_abs=0,
                    absorber_axes=(), sigma_d=None, sigma_b=None, poles=(), pole_masks=None,
                    dtype=torch.float64, device="cuda", absorber_m: int = 2,
                    absorber_R0: float = 1e-15, pmc_axes=()) -> Coeffs3D:
    """stepper_3d.py:150-266: dt = 0.5 dx / sqrt(3).  Boundary modes
    (stepper_3d.py:174-215): CPML (npml > 0, n_abs = 0): per axis an H-side
    profile of length n-1 at offset 1/2 and an E-side profile of length n at
    offset 0 trimmed to [1:-1]; matched-sigma absorber (npml = 0): no CPML
    (cpml None, the absorber path of step_3d) and the conductivity
    sig_bound = max over axes of the absorber_sigma ramps (zero with
    n_abs = 0: a plain PEC box); mixed (both): CPML on the axes outside
    absorber_axes, the identity profile (b = a = 0, k = 1: psi stays zero, a
    plain derivative) on the axes in it, and sig_bound = max of their ramps.
    sig_bound adds to the material sigma_d and sigma_b.  npml / n_abs: an int
    for every face, or per axis an int or (lo, hi) pair.  sigma_d / sigma_b:
    cell-centred (nx, ny, nz) conductivities (stepper_3d.py:217-235); poles:
    Lorentz / Drude (fn, gamma, sigma[, fc]) with cell-centred pole_masks,
    giving per E component c1, c2 of shape (np, 1, 1, 1) and c3 = v3 * (mask
    averaged along the component's axis) (stepper_3d.py:237-266)."""
    from .coeffs import pole_recurrence
    if len(pmc_axes):
        raise NotImplementedError("3D PMC walls are not supported by the HIP stepper")
    npml_ax, nabs_ax = _axis_sides3(npml), _axis_sides3(n_abs)
    has_pml = any(l or h for l, h in npml_ax)
    has_abs = any(l or h for l, h in nabs_ax)
    mixed = has_pml and has_abs
    ident = set(int(a) for a in absorber_axes) if mixed else set()
    if not ident <= {0, 1, 2}:
        raise ValueError(f"absorber_axes must name axes 0, 1, 2, got {tuple(absorber_axes)}")
    dt = 0.5 * dx / math.sqrt(3.0)
    e = (eps if isinstance(eps, torch.Tensor) else torch.ones(nx, ny, nz, dtype=dtype))
    e = e.to(dtype=dtype, device=device)
    dims = (nx, ny, nz)
    ramp = lambda ax: absorber_sigma(dims[ax], nabs_ax[ax], dx, R0=absorber_R0, m=absorber_m, dtype=dtype)
    cpml = shells = sig_bound = None
    if has_pml:
        cpml, shells = {}, {}
        for ax, kh, ke in ((0, "hx", "ex"), (1, "hy", "ey"), (2, "hz", "ez")):
            n = dims[ax]
            if ax in ident:   # one shell over the whole axis: psi' = 0 psi + 0 d, curl d + psi'
                for key, m in ((kh, n - 1), (ke, n - 2)):
                    z = torch.zeros(m, dtype=dtype)
                    cpml[key] = tuple(_bc(v.to(device), ax) for v in (z, z.clone(), torch.ones(m, dtype=dtype)))
                    shells[key] = (m, m)
                continue
            b, a, k = cpml_profiles(n - 1, npml_ax[ax], 0.5, dt, dx, dtype=dtype)
            shells[kh] = shell_bounds(b, a)
            cpml[kh] = tuple(_bc(v.to(device), ax) for v in (b, a, k))
            b, a, k = cpml_profiles(n, npml_ax[ax], 0.0, dt, dx, dtype=dtype)
            b, a, k = b[1:-1].contiguous(), a[1:-1].contiguous(), k[1:-1].contiguous()
            shells[ke] = shell_bounds(b, a)
            cpml[ke] = tuple(_bc(v.to(device), ax) for v in (b, a, k))
        if mixed:
            sig_bound = torch.zeros(nx, ny, nz, dtype=dtype)
            for ax in sorted(ident):
                sig_bound = torch.maximum(sig_bound, _bc(ramp(ax), ax).expand(nx, ny, nz))
    elif has_abs:
        sx, sy, sz = ramp(0), ramp(1), ramp(2)
        sig_bound = torch.maximum(torch.maximum(sx[:, None, None], sy[None, :, None]), sz[None, None, :])
    else:
        sig_bound = torch.zeros(nx, ny, nz, dtype=dtype)
    for name, v in (("sigma_d", sigma_d), ("sigma_b", sigma_b)):
        if v is not None and (not torch.is_tensor(v) or tuple(v.shape) != (nx, ny, nz)):
            raise ValueError(f"{name} must be a ({nx}, {ny}, {nz}) tensor")
    pols = None
    if len(poles):
        if pole_masks is not None and len(pole_masks) != len(poles):
            raise ValueError("pole_masks needs one mask per pole")
        pols = {}
        rec = [pole_recurrence(tuple(p), dt) for p in poles]
        n = len(poles)
        c1 = torch.tensor([r[0] for r in rec], dtype=dtype, device=device).reshape(n, 1, 1, 1)
        c2 = torch.tensor([r[1] for r in rec], dtype=dtype, device=device).reshape(n, 1, 1, 1)
        for key, shape, ax in (("x", (nx - 1, ny, nz), 0), ("y", (nx, ny - 1, nz), 1),
                               ("z", (nx, ny, nz - 1), 2)):
            c3 = torch.zeros(n, *shape, dtype=dtype, device=device)
            for q in range(n):
                m = (pole_masks[q] if pole_masks is not None else
                     torch.ones(nx, ny, nz, dtype=dtype, device=device))
                c3[q] = rec[q][2] * _avg(m.to(dtype=dtype, device=device), ax)
            pols[key] = (c1, c2, c3)
    return Coeffs3D(dt=dt, dx=dx, eps=e, damp=_Damp(e, dt, sigma_d, sigma_b, sig_bound), cpml=cpml, pols=pols,
                    pmc=(), shells=shells)


def has_media(c: Coeffs3D) -> bool:
    """True when the coefficients carry conductivity or ADE poles."""
    return c.pols is not None or bool(getattr(c.damp, "active", False))


def init_3d_state(nx: int, ny: int, nz: int, npml: int = 0, n_poles: int = 0,
                  dtype=torch.float64, device="cuda") -> State3D:
    """stepper_3d.py:269-292 (pole stacks (n_poles, *E component shape))."""
    z = lambda *s: torch.zeros(*s, dtype=dtype, device=device)
    psi = {}
    if npml > 0:
        psi = dict(
            psi_Ez_y=z(nx, ny - 1, nz - 1), psi_Ey_z=z(nx, ny - 1, nz - 1),
            psi_Ex_z=z(nx - 1, ny, nz - 1), psi_Ez_x=z(nx - 1, ny, nz - 1),
            psi_Ey_x=z(nx - 1, ny - 1, nz), psi_Ex_y=z(nx - 1, ny - 1, nz),
            psi_Hz_y=z(nx - 1, ny - 2, nz), psi_Hy_z=z(nx - 1, ny, nz - 2),
            psi_Hx_z=z(nx, ny - 1, nz - 2), psi_Hz_x=z(nx - 2, ny - 1, nz),
            psi_Hy_x=z(nx - 2, ny, nz - 1), psi_Hx_y=z(nx, ny - 2, nz - 1))
    if n_poles:
        psi.update(Px_cur=z(n_poles, nx - 1, ny, nz), Px_prev=z(n_poles, nx - 1, ny, nz),
                   Py_cur=z(n_poles, nx, ny - 1, nz), Py_prev=z(n_poles, nx, ny - 1, nz),
                   Pz_cur=z(n_poles, nx, ny, nz - 1), Pz_prev=z(n_poles, nx, ny, nz - 1))
    return State3D(z(nx - 1, ny, nz), z(nx, ny - 1, nz), z(nx, ny, nz - 1),
                   z(nx, ny - 1, nz - 1), z(nx - 1, ny, nz - 1), z(nx - 1, ny - 1, nz), **psi)


POLE_NAMES = ("Px_cur", "Px_prev", "Py_cur", "Py_prev", "Pz_cur", "Pz_prev")


def state_names(state: State3D):
    """Field names a run advances: components, psi (a CPML state) and, with
    poles, the stacks."""
    return (COMPS + (PSI_NAMES if state.psi_Ez_y is not None else ()) +
            (POLE_NAMES if state.Px_cur is not None else ()))


class _OneStep:
    def __init__(self, amp, src):
        self._amp, self.src = float(amp), src

    def amplitude(self, t):
        return self._amp


def step_3d(state: State3D, c: Coeffs3D, src_amp: float = 0.0, src=None) -> State3D:
    """One leapfrog step (stepper_3d.py:303-445) on the HIP engine.  Unlike the
    reference's point-source path, which writes into the new state's tensor in
    place (stepper_3d.py:476-481), the input state is never modified."""
    from .engine3d import run_fields
    work = {f: getattr(state, f).clone() for f in state_names(state)}
    run_fields(work, c, 1, _OneStep(src_amp, src) if src is not None else None, [], lean=False)
    return state._replace(**work)
