"""Lean in-place forward 3D stepper (meep_torch/core/stepper_3d_lean.py).

Lean3D keeps the reference attributes: F (fields, updated in place), psi as
(lo, hi) shell views, inv_eps, and with ADE poles pol / P.  The E update and the source sign follow the
lean form (E += curl * inv_eps * dt; sources subtract), which rounds
differently from step_3d.  ``compile`` is accepted and ignored: the step is
already one pair of fused HIP kernels."""
from dataclasses import dataclass, field

import torch

from .engine3d import run_fields
from .stepper_3d import build_3d_coeffs

__all__ = ["build_lean", "run_3d_lean", "step_lean", "Lean3D"]

_PSI_AXIS = {"psi_Ez_y": 1, "psi_Ey_z": 2, "psi_Ex_z": 2, "psi_Ez_x": 0, "psi_Ey_x": 0,
             "psi_Ex_y": 1, "psi_Hz_y": 1, "psi_Hy_z": 2, "psi_Hx_z": 2, "psi_Hz_x": 0,
             "psi_Hy_x": 0, "psi_Hx_y": 1}


@dataclass
class Lean3D:
    dt: float
    dx: float
    nx: int
    ny: int
    nz: int
    npml: int
    inv_eps: dict
    coeffs: object
    F: dict = field(default_factory=dict)
    psi: dict = field(default_factory=dict)
    full_psi: dict = field(default_factory=dict)
    # ADE poles (stepper_3d_lean.py:59-64): pol[key] = (c1, c2, c3), P[key] =
    # [P(n+1), P(n)] (the state's P?_cur / P?_prev stacks, advanced in place)
    pol: dict = field(default_factory=dict)
    P: dict = field(default_factory=dict)


def build_lean(nx, ny, nz, dx, eps, npml, poles=(), pole_masks=None, dtype=torch.float32,
               device="cuda") -> Lean3D:
    """stepper_3d_lean.py:90-165; poles with cell-centred masks as in
    build_3d_coeffs (the same c1, c2 scalars and Yee-averaged c3 stacks)."""
    c = build_3d_coeffs(nx, ny, nz, dx, eps=eps, npml=npml, poles=poles, pole_masks=pole_masks, dtype=dtype,
                        device=device)
    e = c.eps
    inv = {k: 1.0 / (0.5 * (e.narrow(ax, 0, e.shape[ax] - 1) + e.narrow(ax, 1, e.shape[ax] - 1)))
           for k, ax in (("x", 0), ("y", 1), ("z", 2))}
    z = lambda *s: torch.zeros(*s, dtype=dtype, device=device)
    F = dict(Ex=z(nx - 1, ny, nz), Ey=z(nx, ny - 1, nz), Ez=z(nx, ny, nz - 1),
             Hx=z(nx, ny - 1, nz - 1), Hy=z(nx - 1, ny, nz - 1), Hz=z(nx - 1, ny - 1, nz))
    shapes = {"psi_Ez_y": (nx, ny - 1, nz - 1), "psi_Ey_z": (nx, ny - 1, nz - 1),
              "psi_Ex_z": (nx - 1, ny, nz - 1), "psi_Ez_x": (nx - 1, ny, nz - 1),
              "psi_Ey_x": (nx - 1, ny - 1, nz), "psi_Ex_y": (nx - 1, ny - 1, nz),
              "psi_Hz_y": (nx - 1, ny - 2, nz), "psi_Hy_z": (nx - 1, ny, nz - 2),
              "psi_Hx_z": (nx, ny - 1, nz - 2), "psi_Hz_x": (nx - 2, ny - 1, nz),
              "psi_Hy_x": (nx - 2, ny, nz - 1), "psi_Hx_y": (nx, ny - 2, nz - 1)}
    full = {n: z(*s) for n, s in shapes.items()}
    psi = {}
    for n, t in full.items():
        ax = _PSI_AXIS[n]
        side = "e" if n.startswith("psi_H") else "h"
        key = side + "xyz"[ax]
        lo, hi = c.shells[key]
        m = t.shape[ax]
        psi[n[4:]] = (t.narrow(ax, 0, max(lo, 1)), t.narrow(ax, min(hi, m - 1), m - min(hi, m - 1)))
    lean = Lean3D(dt=c.dt, dx=dx, nx=nx, ny=ny, nz=nz, npml=npml, inv_eps=inv, coeffs=c,
                  F=F, psi=psi, full_psi=full)
    if c.pols is not None:
        n = len(poles)
        for key in "xyz":
            lean.pol[key] = c.pols[key]
            shape = (n,) + tuple(F["E" + key].shape)
            lean.P[key] = [z(*shape), z(*shape)]
    return lean


class _StateView:
    """Duck-typed State3D view over the in-place field dict (stepper_3d_lean.py:299-304)."""

    def __init__(self, F):
        self.__dict__.update(F)


class _OneStep:
    def __init__(self, amp, src):
        self._amp, self.src = float(amp), src

    def amplitude(self, t):
        return self._amp


def _fields(c: Lean3D):
    d = dict(c.F)
    d.update(c.full_psi)
    for key, (p_new, p_old) in c.P.items():
        d[f"P{key}_cur"], d[f"P{key}_prev"] = p_new, p_old
    return d


@torch.no_grad()
def step_lean(c: Lean3D, src_amp=0.0, src=None):
    """One in-place lean step (stepper_3d_lean.py:234-296)."""
    run_fields(_fields(c), c.coeffs, 1, _OneStep(src_amp, src) if src is not None else None, [],
               lean=True)


@torch.no_grad()
def run_3d_lean(c: Lean3D, n_steps: int, *, source=None, monitors=(), compile: bool = False):
    """stepper_3d_lean.py:307-349: fields advance in place; returns
    (state view, accumulators)."""
    outs = run_fields(_fields(c), c.coeffs, n_steps, source, list(monitors), lean=True)
    return _StateView(c.F), outs
