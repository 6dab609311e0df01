"""One-way TFSF eigenmode sources (meep_torch/ports/mode_source.py:21-128).

A port's guided mode is launched by an electric sheet (the engine's soft
source mask) and a magnetic sheet half a cell upstream on the launch side
(the h_masks / h_amplitude currents of the B phase), both SRC_OFFSET cells
upstream of the port's monitor plane.  The sheets carry the mode's
transverse H and E profiles, and the magnetic sheet fires later by the
half-cell phase delay neff dx / 2, so the two radiate in step downstream
and cancel upstream.  Profiles come from the 1D mode solve of the eps line
at the monitor plane and are not differentiated through."""
import math

import torch

from .modes1d import solve_1d
from .port import SRC_OFFSET, TE_SRC_OFFSET

# relative signs of the magnetic sheets (ports/mode_source.py:13-14)
H_SIGN_TM = -1.0
H_SIGN_TE = 1.0


def _line(t, port):
    return t[port.index, port.lo:port.hi] if port.axis == "x" else t[port.lo:port.hi, port.index]


def _poles_on_line(c3, port):
    return c3 is not None and c3.numel() > 0 and bool((c3[:, port.index, port.lo:port.hi] != 0).any()
                                                       if port.axis == "x" else
                                                       (c3[:, port.lo:port.hi, port.index] != 0).any())


def _guided(modes, mode, clad_eps):
    neff = float(modes.neff[mode])
    clad = math.sqrt(float(clad_eps))
    if neff <= clad + 1e-9:
        raise ValueError(f"no guided mode at port (neff={neff:.4f} <= cladding {clad:.4f})")
    return neff


class ModeSourceTM:
    """TM (Ez, Hx, Hy) mode source on an x- or y-normal port
    (ports/mode_source.py:21-81).  bc='pmc_lo' solves the upper half of an
    even mode for a half-domain port whose low end is a mirror plane."""

    def __init__(self, grid, coeffs, port, profile, fcen, mode=0, bc="pec"):
        self.profile, self.dt, self.d = profile, grid.dt, port.direction
        dx = grid.dx
        eps_line = _line(coeffs.eps_inf, port).detach()
        if _poles_on_line(getattr(coeffs, "c3", None), port):
            raise ValueError("dispersive media on port line unsupported")
        modes = solve_1d(eps_line, dx, fcen, "tm", k=mode + 1, bc=bc)
        clad = eps_line[-1] if bc == "pmc_lo" else torch.maximum(eps_line[0], eps_line[-1])
        neff = self.neff = _guided(modes, mode, clad)
        prof = modes.profiles[mode].to(device=grid.device, dtype=grid.dtype)
        self.delay = neff * dx / 2.0
        j = port.index - self.d * SRC_OFFSET
        back = j - 1 if self.d > 0 else j          # the magnetic sheet's row, launch side
        z = lambda *s: torch.zeros(*s, dtype=grid.dtype, device=grid.device)
        self.mask = z(grid.nx, grid.ny)
        if port.axis == "x":
            self.mask[j, port.lo:port.hi] = neff * prof / dx
            hy = z(grid.nx - 1, grid.ny)
            hy[back, port.lo:port.hi] = self.d * prof / dx
            self.h_masks = {"src_hy_mask": hy}
        else:                                      # y-normal: pair Ez / Hx, S_y = +Ez Hx
            self.mask[port.lo:port.hi, j] = -neff * prof / dx
            hx = z(grid.nx, grid.ny - 1)
            hx[port.lo:port.hi, back] = self.d * prof / dx
            self.h_masks = {"src_hx_mask": hx}

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude((t_idx + 0.5) * self.dt)

    def h_amplitude(self, t_idx: int) -> float:
        return H_SIGN_TM * self.profile.amplitude(t_idx * self.dt - self.delay)


class ModeSourceTE:
    """TE (Hz, Ex, Ey) mode source on an x-normal port
    (ports/mode_source.py:84-128): a Dy sheet carrying Hz_m and an Hz sheet
    carrying Ey_m = neff Hz_m / eps; the TE stagger fires the electric sheet
    at t dt and the magnetic one half a step later."""

    def __init__(self, grid, coeffs, port, profile, fcen, mode=0):
        if port.axis != "x":
            raise NotImplementedError("TE y-normal not supported (FluxMonitorTE is x-normal only)")
        self.profile, self.dt, self.d = profile, grid.dt, port.direction
        dx = grid.dx
        eps_line = coeffs.eps_y[port.index, port.lo:port.hi].detach()
        if _poles_on_line(getattr(coeffs, "c3y", None), port):
            raise ValueError("dispersive media on port line unsupported")
        modes = solve_1d(eps_line, dx, fcen, "te", k=mode + 1)
        neff = self.neff = _guided(modes, mode, torch.maximum(eps_line[0], eps_line[-1]))
        hz_m = modes.profiles[mode].to(device=grid.device, dtype=grid.dtype)
        ey_m = (neff * hz_m / eps_line.to(device=grid.device)).to(grid.dtype)
        self.delay = neff * dx / 2.0
        j = port.index - self.d * TE_SRC_OFFSET
        back = j - 1 if self.d > 0 else j
        self.mask = torch.zeros(grid.nx, grid.ny - 1, dtype=grid.dtype, device=grid.device)
        self.mask[j, port.lo:port.hi] = hz_m / dx
        hz = torch.zeros(grid.nx - 1, grid.ny - 1, dtype=grid.dtype, device=grid.device)
        hz[back, port.lo:port.hi] = self.d * ey_m / dx
        self.h_masks = {"src_hz_mask": hz}

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude(t_idx * self.dt)

    def h_amplitude(self, t_idx: int) -> float:
        return H_SIGN_TE * self.profile.amplitude((t_idx + 0.5) * self.dt - self.delay)
