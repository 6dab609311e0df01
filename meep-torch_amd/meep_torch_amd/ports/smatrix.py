"""Sequential per-port S-matrix driver (meep_torch/ports/smatrix.py:38-102).

For every (port j, mode n) column one engine run launches that mode from
port j's source sheets; every port's flux line is decomposed into modal
amplitudes, and S[:, (i, m), (j, n)] = b_out / a_in with a_in the launched
wave at port j (a+ for a +direction port, a- otherwise) and b_out the wave
leaving port i (a- for a +direction port, a+ otherwise).  The phase
reference plane is each port's monitor line.  Forward only."""
from typing import NamedTuple

import torch

from ..core.coeffs import build_te_coeffs, build_tm_coeffs
from ..core.monitors import FluxMonitorTE, FluxMonitorTM
from ..core.run import run_te, run_tm, run_until_decayed
from ..core.sources import GaussianPulse
from ..core.stepper_te import init_te_state, step_te
from ..core.stepper_tm import init_tm_state, step_tm
from .mode_monitor import decompose
from .mode_source import ModeSourceTE, ModeSourceTM
from .port import validate_ports


class SMatrix(NamedTuple):
    s: torch.Tensor          # (nfreq, P, P) complex128, P = total port modes
    freqs: torch.Tensor
    index: list              # column / row -> (port, mode)


def compute_smatrix(eps, grid, ports, *, pol, fcen, fwidth, freqs, n_steps, npml, checkpoint_every=0,
                    decay_probe=None, decay_ratio=1e-6, n_decay=50):
    """eps: cell-centred (nx, ny) map; a Gaussian pulse (fcen, fwidth) drives
    every run for n_steps steps, or, with decay_probe = (ix, iy), until the
    squared field (Ez for TM, Hz for TE) at that cell stays below
    decay_ratio times its peak for n_decay consecutive steps, n_steps at
    most (core/run.py:run_until_decayed)."""
    validate_ports(ports, grid, npml)
    if pol == "tm":
        coeffs = build_tm_coeffs(grid, eps_inf=eps, npml=npml)
        run, init, src_cls = run_tm, init_tm_state, ModeSourceTM
        monitor = lambda p: FluxMonitorTM(grid, freqs, x_normal=p.axis == "x", idx=p.index, lo=p.lo, hi=p.hi)
        line = lambda p: eps[p.index, p.lo:p.hi] if p.axis == "x" else eps[p.lo:p.hi, p.index]
        sign = lambda p: -1.0 if p.axis == "x" else 1.0
    elif pol == "te":
        if any(p.axis != "x" for p in ports):
            raise NotImplementedError("TE y-normal not supported (FluxMonitorTE is x-normal only)")
        coeffs = build_te_coeffs(grid, eps_inf=eps, npml=npml)
        run, init, src_cls = run_te, init_te_state, ModeSourceTE
        monitor = lambda p: FluxMonitorTE(grid, freqs, ix=p.index, lo=p.lo, hi=p.hi)
        # eps on the Ey nodes: the half-y average of the cell-centred rows
        line = lambda p: 0.5 * (eps[p.index, p.lo:p.hi] + eps[p.index, p.lo + 1:p.hi + 1])
        sign = lambda p: -1.0
    else:
        raise ValueError(f"pol must be 'tm' or 'te', got {pol!r}")
    index = [(j, n) for j, p in enumerate(ports) for n in range(p.n_modes)]
    s = torch.zeros(len(freqs), len(index), len(index), dtype=torch.complex128)
    for col, (j, n) in enumerate(index):
        src = src_cls(grid, coeffs, ports[j], GaussianPulse(fcen, fwidth), fcen, mode=n)
        mons = [monitor(p) for p in ports]
        if decay_probe is not None:
            ix, iy = decay_probe
            f = "Ez" if pol == "tm" else "Hz"
            _, accs = run_until_decayed(step_tm if pol == "tm" else step_te, init(grid), coeffs, src, mons,
                                        lambda st: getattr(st, f)[ix, iy] * getattr(st, f)[ix, iy],
                                        decay_ratio=decay_ratio, n_decay=n_decay, max_steps=n_steps)
        else:
            _, accs = run(init(grid), coeffs, n_steps, source=src, monitors=mons, checkpoint_every=checkpoint_every)
        amps = [decompose(acc, line(p), grid.dx, freqs, pol, p.n_modes, axis_sign=sign(p))
                for p, acc in zip(ports, accs)]
        a_p, a_m = amps[j]
        a_in = (a_p if ports[j].direction > 0 else a_m)[:, n]
        for row, (i, m) in enumerate(index):
            b_p, b_m = amps[i]
            s[:, row, col] = ((b_m if ports[i].direction > 0 else b_p)[:, m] / a_in).cpu()
    return SMatrix(s=s, freqs=torch.tensor([float(f) for f in freqs], dtype=torch.float64), index=index)
