"""2D TE state and single step (meep_torch/core/stepper_te.py:11-49, 178-190)."""
from typing import NamedTuple

import torch

from .grid import Grid


class TEState(NamedTuple):
    Hz: torch.Tensor       # (nx-1, ny-1) at n
    Ex: torch.Tensor       # (nx-1, ny)   at n-1/2
    Ey: torch.Tensor       # (nx, ny-1)   at n-1/2
    Dx: torch.Tensor       # (nx-1, ny)
    Dy: torch.Tensor       # (nx, ny-1)
    Px_cur: torch.Tensor   # (np, nx-1, ny)
    Px_prev: torch.Tensor
    Py_cur: torch.Tensor   # (np, nx, ny-1)
    Py_prev: torch.Tensor
    psi_Eyx: torch.Tensor  # (nx-1, ny-1)
    psi_Exy: torch.Tensor  # (nx-1, ny-1)
    psi_Hzy: torch.Tensor  # (nx-1, ny-2)
    psi_Hzx: torch.Tensor  # (nx-2, ny-1)


FIELDS = ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx")


def init_te_state(grid: Grid, n_poles: int = 0, n_poles_y: int | None = None) -> TEState:
    """stepper_te.py:27-49."""
    npx = int(n_poles)
    npy = npx if n_poles_y is None else int(n_poles_y)
    nx, ny = grid.nx, grid.ny
    return TEState(
        Hz=grid.zeros(nx - 1, ny - 1), Ex=grid.zeros(nx - 1, ny), Ey=grid.zeros(nx, ny - 1),
        Dx=grid.zeros(nx - 1, ny), Dy=grid.zeros(nx, ny - 1),
        Px_cur=grid.zeros(npx, nx - 1, ny), Px_prev=grid.zeros(npx, nx - 1, ny),
        Py_cur=grid.zeros(npy, nx, ny - 1), Py_prev=grid.zeros(npy, nx, ny - 1),
        psi_Eyx=grid.zeros(nx - 1, ny - 1), psi_Exy=grid.zeros(nx - 1, ny - 1),
        psi_Hzy=grid.zeros(nx - 1, ny - 2), psi_Hzx=grid.zeros(nx - 2, ny - 1))


def fields_of(state: TEState):
    """State fields the engine advances: FIELDS plus the pole stacks the
    state carries."""
    extra = (("Px_cur", "Px_prev") if state.Px_cur.shape[0] else ()) + \
        (("Py_cur", "Py_prev") if state.Py_cur.shape[0] else ())
    return FIELDS + extra


def _check(state: TEState, m=None):
    if m is not None and (state.Px_cur.shape[0] != m.c1x.shape[0] or state.Py_cur.shape[0] != m.c1y.shape[0]):
        raise ValueError(f"state has {state.Px_cur.shape[0]} / {state.Py_cur.shape[0]} pole stacks, "
                         f"coefficients {m.c1x.shape[0]} / {m.c1y.shape[0]} poles")


def step_te(state: TEState, m, src_amp=0.0, src_mask_ey: torch.Tensor | None = None,
            src_h_amp: float = 0.0, src_hz_mask: torch.Tensor | None = None) -> TEState:
    """One TE step entered at the D half (stepper_te.py:178-190) on the engine."""
    from .run import _Fixed
    from . import engine2d
    _check(state, m)
    work = {f: getattr(state, f).clone() for f in fields_of(state)}
    plan = None
    if src_mask_ey is not None or src_hz_mask is not None:
        nx, ny = work["Ey"].shape[0], work["Ex"].shape[1]
        plan = engine2d.source_plan(_Fixed(src_amp, src_mask_ey, src_h_amp, {"src_hz_mask": src_hz_mask}),
                                    work["Dy"].shape, work["Dy"].device, work["Dy"].dtype,
                                    engine2d.h_shapes("te", nx, ny))
    engine2d.run_chunk("te", work, m, 0, 1, plan, [])
    return state._replace(**work)
