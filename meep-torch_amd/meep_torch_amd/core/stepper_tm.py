"""2D TM state and single step (meep_torch/core/stepper_tm.py:12-71, 234-254).

TMState keeps the reference fields and shapes.  step_tm advances one step on
the HIP engine (functional: the input state is not modified)."""
from typing import NamedTuple

import torch

from .coeffs import TMCoeffs
from .grid import Grid


class TMState(NamedTuple):
    Ez: torch.Tensor       # (nx, ny)   E_z at time n
    Hx: torch.Tensor       # (nx, ny-1) at n-1/2
    Hy: torch.Tensor       # (nx-1, ny) at n-1/2
    Dz: torch.Tensor       # (nx, ny)
    P_cur: torch.Tensor    # (n_poles, nx, ny)
    P_prev: torch.Tensor
    psi_Ezy: torch.Tensor  # (nx, ny-1)
    psi_Ezx: torch.Tensor  # (nx-1, ny)
    psi_Hyx: torch.Tensor  # (nx-2, ny_e)
    psi_Hxy: torch.Tensor  # (nx-2, ny_e)
    Bx: torch.Tensor | None = None
    By: torch.Tensor | None = None
    Mx_cur: torch.Tensor | None = None
    Mx_prev: torch.Tensor | None = None
    My_cur: torch.Tensor | None = None
    My_prev: torch.Tensor | None = None
    Hx0: torch.Tensor | None = None
    Hy0: torch.Tensor | None = None


FIELDS = ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy")
MAG_FIELDS = ("Bx", "By", "Mx_cur", "Mx_prev", "My_cur", "My_prev")


def init_tm_state(grid: Grid, n_poles: int = 0, y_pmc: bool = False, n_mpoles: int = 0,
                  y_mirror: int = 0, x_mirror: int = 0, y_rot2: int = 0, rot4: int = 0) -> TMState:
    """stepper_tm.py:37-71 (psi_H rows: ny-2 with PEC walls, ny under y_pmc,
    ny-1 under a low-y cut; psi_H columns nx-2, nx-1 under a low-x cut; the
    on-plane rows Hx0 (nx,) under an odd y mirror, Rotate2 or Rotate4 and
    Hy0 (ny,) under an odd x mirror or Rotate4; with n_mpoles the B fields
    and the magnetic pole stacks)."""
    nx, ny = grid.nx, grid.ny
    ny_e = ny if y_pmc else (ny - 1 if (y_mirror or y_rot2 or rot4) else ny - 2)
    nx_e = nx - 1 if (x_mirror or rot4) else nx - 2
    mag = {}
    if n_mpoles:
        mag = dict(Bx=grid.zeros(nx, ny - 1), By=grid.zeros(nx - 1, ny),
                   Mx_cur=grid.zeros(n_mpoles, nx, ny - 1), Mx_prev=grid.zeros(n_mpoles, nx, ny - 1),
                   My_cur=grid.zeros(n_mpoles, nx - 1, ny), My_prev=grid.zeros(n_mpoles, nx - 1, ny))
    return TMState(
        Ez=grid.zeros(nx, ny), Hx=grid.zeros(nx, ny - 1), Hy=grid.zeros(nx - 1, ny),
        Dz=grid.zeros(nx, ny), P_cur=grid.zeros(n_poles, nx, ny), P_prev=grid.zeros(n_poles, nx, ny),
        psi_Ezy=grid.zeros(nx, ny - 1), psi_Ezx=grid.zeros(nx - 1, ny),
        psi_Hyx=grid.zeros(nx_e, ny_e), psi_Hxy=grid.zeros(nx_e, ny_e),
        Hx0=grid.zeros(nx) if (y_mirror == -1 or y_rot2 or rot4) else None,
        Hy0=grid.zeros(ny) if (x_mirror == -1 or rot4) else None, **mag)


def fields_of(state: TMState):
    """State fields the engine advances: FIELDS plus the pole stacks when the
    state carries poles, and B with the magnetic pole stacks."""
    return (FIELDS + (("P_cur", "P_prev") if state.P_cur.shape[0] else ()) +
            (MAG_FIELDS if state.Bx is not None else ()) + (("Hx0",) if state.Hx0 is not None else ()) +
            (("Hy0",) if state.Hy0 is not None else ()))


def _check(state: TMState, m: TMCoeffs):
    want_hx0 = m.y_mirror_lo == -1 or bool(m.y_rot2_lo) or bool(m.rot4_lo)
    want_hy0 = m.x_mirror_lo == -1 or bool(m.rot4_lo)
    if (state.Hx0 is not None) != want_hx0 or (state.Hy0 is not None) != want_hy0:
        raise ValueError("the state's on-plane H rows do not match the coefficients' symmetry cuts "
                         "(init_tm_state with the same y_mirror / x_mirror / y_rot2 / rot4)")
    nm = 0 if m.mc1x is None else int(m.mc1x.shape[0])
    if (state.Bx is not None) != bool(nm) or (nm and int(state.Mx_cur.shape[0]) != nm):
        raise ValueError(f"magnetic poles: the coefficients have {nm}, the state "
                         f"{0 if state.Mx_cur is None else int(state.Mx_cur.shape[0])} (init_tm_state(n_mpoles=...))")
    if state.P_cur.shape[0] != m.c1.shape[0]:
        raise ValueError(f"state has {state.P_cur.shape[0]} pole stacks, coefficients {m.c1.shape[0]} poles")
    nx, ny = state.Ez.shape
    ny_e = ny if m.y_pmc else (ny - 1 if (m.y_mirror_lo or m.y_rot2_lo or m.rot4_lo) else ny - 2)
    nx_e = nx - 1 if (m.x_mirror_lo or m.rot4_lo) else nx - 2
    if tuple(state.psi_Hxy.shape) != (nx_e, ny_e):
        raise ValueError(f"psi_Hxy shape {tuple(state.psi_Hxy.shape)} does not match y_pmc={m.y_pmc} "
                         f"and the symmetry cuts")


def step_tm(state: TMState, m: TMCoeffs, src_amp=0.0, src_mask: torch.Tensor | None = None,
            src_h_amp: float = 0.0, src_hy_mask: torch.Tensor | None = None,
            src_hx_mask: torch.Tensor | None = None) -> TMState:
    """One TM step in meep's phase order (B, source(B), eh(H), D, source(D),
    pols, eh(E)) on the HIP engine."""
    from .run import _Fixed
    from . import engine2d
    _check(state, m)
    work = {f: getattr(state, f).clone() for f in fields_of(state)}
    plan = None
    h = {"src_hy_mask": src_hy_mask, "src_hx_mask": src_hx_mask}
    if src_mask is not None or any(v is not None for v in h.values()):
        nx, ny = work["Ez"].shape
        plan = engine2d.source_plan(_Fixed(src_amp, src_mask, src_h_amp, h), work["Dz"].shape,
                                    work["Dz"].device, work["Dz"].dtype, engine2d.h_shapes("tm", nx, ny))
    engine2d.run_chunk("tm", work, m, 0, 1, plan, [])
    return state._replace(**work)
