"""2D monitors (meep_torch/core/monitors.py:11-155).

Each class keeps the reference constructor, ``reduce`` attribute, result shape
and ``flux()`` post-processing.  Sampling happens inside the engine: a monitor
describes itself as up to four strided field views ("terms"), each sampled
after every step and folded into dt*[cos(w t) v, -sin(w t) v] accumulators
(monitors.py:84-95) by the DFT kernels.  Term times: E at (t+1)dt and H at
(t+1/2)dt for TM, E at (t+1/2)dt and Hz at (t+1)dt for TE (monitors.py:84-85,
116-118, 145-147); DFTFieldsMonitor uses the TM timing for every run
(monitors.py:42-43)."""
import math
from typing import Sequence

import torch

from .grid import Grid


def _line(dim: int, lo: int, hi: int):
    r = range(dim)[slice(lo, hi)]
    return r.start, len(r)


TM_IDS = {"Ez": 0, "Hx": 1, "Hy": 2, "Dz": 3}
TE_IDS = {"Hz": 0, "Ex": 1, "Ey": 2, "Dx": 3, "Dy": 4}


def _ids(f):
    return TM_IDS if "Ez" in f else TE_IDS


class PointProbe:
    reduce = "stack"

    def __init__(self, ix: int, iy: int):
        self.ix, self.iy = ix, iy

    def engine_terms(self, f):
        nx, ny = f["Ez"].shape
        sem = (0, 4, self.ix % nx, self.iy % ny)
        return "probe", [(f["Ez"][self.ix, self.iy].reshape(1, 1), None, 0, sem)]


class DFTFieldsMonitor:
    """Running DFT of one full-grid component; result (nfreq, 2, *shape)."""

    reduce = "sum"

    def __init__(self, grid: Grid, freqs: Sequence[float], component: str = "Ez"):
        self.component = component
        self.dt = grid.dt
        self.freqs = torch.tensor(freqs, dtype=grid.dtype, device=grid.device)

    def engine_terms(self, f):
        F = f[self.component]
        half = 0 if self.component.startswith("E") else 1
        fid = _ids(f).get(self.component)
        sem = (fid, 3, F.shape[0], F.shape[1]) if fid is not None else None
        return "dft", [(F.reshape(1, -1), None, half, sem)]

    def out_shape(self, f):
        return (self.freqs.numel(), 2) + tuple(f[self.component].shape)


class FluxMonitorTM:
    reduce = "sum"

    def __init__(self, grid: Grid, freqs: Sequence[float], *, x_normal: bool, idx: int,
                 lo: int, hi: int):
        self.x_normal = x_normal
        self.idx, self.lo, self.hi = idx, lo, hi
        self.dt = grid.dt
        self.freqs = torch.tensor(freqs, dtype=grid.dtype, device=grid.device)
        if x_normal and not (1 <= idx < grid.nx - 1):
            raise ValueError(f"x-normal monitor ix={idx} must be interior")
        if not x_normal and not (1 <= idx < grid.ny - 1):
            raise ValueError(f"y-normal monitor iy={idx} must be interior")

    def engine_terms(self, f):
        i, lo, hi = self.idx, self.lo, self.hi
        nx, ny = f["Ez"].shape
        if self.x_normal:
            E = f["Ez"][i, lo:hi]
            Ha, Hb = f["Hy"][i - 1, lo:hi], f["Hy"][i, lo:hi]
            l0 = _line(ny, lo, hi)[0]
            return "dft", [(E, None, 0, (0, 1, i, l0)), (Ha, Hb, 1, (2, 1, i, l0))]
        E = f["Ez"][lo:hi, i]
        Ha, Hb = f["Hx"][lo:hi, i - 1], f["Hx"][lo:hi, i]
        l0 = _line(nx, lo, hi)[0]
        return "dft", [(E, None, 0, (0, 2, i, l0)), (Ha, Hb, 1, (1, 2, i, l0))]

    def out_shape(self, f):
        n = (f["Ez"].shape[1] if self.x_normal else f["Ez"].shape[0])
        return (self.freqs.numel(), 4, _line(n, self.lo, self.hi)[1])

    def flux(self, acc: torch.Tensor, dx: float) -> torch.Tensor:
        ReE, ImE, ReH, ImH = acc.unbind(1)
        s = (ReE * ReH + ImE * ImH).sum(-1) * dx
        return -0.5 * s if self.x_normal else 0.5 * s


class FluxMonitorTE:
    reduce = "sum"

    def __init__(self, grid: Grid, freqs, *, ix: int, lo: int, hi: int):
        if not (1 <= ix < grid.nx - 1):
            raise ValueError(f"ix={ix} must be interior")
        self.ix, self.lo, self.hi = ix, lo, hi
        self.dt = grid.dt
        self.freqs = torch.tensor(freqs, dtype=grid.dtype, device=grid.device)

    def engine_terms(self, f):
        i, lo, hi = self.ix, self.lo, self.hi
        l0 = _line(f["Ey"].shape[1], lo, hi)[0]
        return "dft", [(f["Ey"][i, lo:hi], None, 1, (2, 1, i, l0)),
                       (f["Hz"][i - 1, lo:hi], f["Hz"][i, lo:hi], 0, (0, 1, i, l0))]

    def out_shape(self, f):
        return (self.freqs.numel(), 4, _line(f["Ey"].shape[1], self.lo, self.hi)[1])

    def flux(self, acc: torch.Tensor, dx: float) -> torch.Tensor:
        ReE, ImE, ReH, ImH = acc.unbind(1)
        return 0.5 * (ReE * ReH + ImE * ImH).sum(-1) * dx


class FluxMonitorTEY:
    reduce = "sum"

    def __init__(self, grid: Grid, freqs, *, iy: int, lo: int, hi: int):
        if not (1 <= iy < grid.ny - 1):
            raise ValueError(f"iy={iy} must be interior")
        self.iy, self.lo, self.hi = iy, lo, hi
        self.dt = grid.dt
        self.freqs = torch.tensor(freqs, dtype=grid.dtype, device=grid.device)

    def engine_terms(self, f):
        j, lo, hi = self.iy, self.lo, self.hi
        l0 = _line(f["Ex"].shape[0], lo, hi)[0]
        return "dft", [(f["Ex"][lo:hi, j], None, 1, (1, 2, j, l0)),
                       (f["Hz"][lo:hi, j - 1], f["Hz"][lo:hi, j], 0, (0, 2, j, l0))]

    def out_shape(self, f):
        return (self.freqs.numel(), 4, _line(f["Ex"].shape[0], self.lo, self.hi)[1])

    def flux(self, acc: torch.Tensor, dx: float) -> torch.Tensor:
        ReE, ImE, ReH, ImH = acc.unbind(1)
        return -0.5 * (ReE * ReH + ImE * ImH).sum(-1) * dx


def omega_of(mon) -> torch.Tensor:
    """2*pi*f in the grid dtype, formed as the reference does (monitors.py:86)."""
    return (2 * math.pi * mon.freqs).contiguous()
