"""S-matrix of a 2D device (the interface of meep_torch/ports/smatrix.py:38-102).

A column of S -- the response of every port to one mode n launched from one
port j -- is a simulation of its own that differs from the other columns
only in its source sheets.  On this engine the TM columns therefore run
together: each column is a variant of ONE BatchedTM (persistent launch over
the tiles of all columns, the TFSF magnetic sheets injected in-kernel), and
the spectra of every port line come back from that single run.  TE columns,
and runs stopped by a field-decay probe, step one column at a time.

Reading a column: every port line's spectra are split into forward and
backward modal amplitudes (decompose); the wave a port launches is the one
travelling along its direction, the wave it receives the opposite one, and
S[f, (i, m), (j, n)] = received_i,m / launched_j,n with each port's monitor
line as its phase reference plane.  Forward only (sources are built from
detached mode profiles)."""
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


def _port_line_monitor(grid, freqs, pol, port):
    if pol == "tm":
        return FluxMonitorTM(grid, freqs, x_normal=port.axis == "x", idx=port.index, lo=port.lo, hi=port.hi)
    return FluxMonitorTE(grid, freqs, ix=port.index, lo=port.lo, hi=port.hi)


def _port_eps_line(eps, pol, port):
    """eps on the port line's E nodes: the cell-centred values for TM (Ez on
    cell centres), the half-y average for TE (Ey nodes)."""
    if pol == "tm":
        return eps[port.index, port.lo:port.hi] if port.axis == "x" else eps[port.lo:port.hi, port.index]
    return 0.5 * (eps[port.index, port.lo:port.hi] + eps[port.index, port.lo + 1:port.hi + 1])


def _flux_sign(pol, port):
    """Sign of the stored H against E for a forward TM wave: Poynting flux is
    -Ez Hy along x and +Ez Hx along y (decompose ignores it for TE)."""
    return 1.0 if (pol == "tm" and port.axis == "y") else -1.0


def _column_accs_batched(coeffs, eps, sources, monitor_sets, n_steps):
    """Spectra of every column from one batched launch (all columns share
    the grid, eps and CPML; only the source sheets differ)."""
    from ..batch import BatchedTM
    probs = [(coeffs, src, mons) for src, mons in zip(sources, monitor_sets)]
    bt = BatchedTM(probs, n_steps, checkpoint_every=0, adjoint=False)
    out = bt.forward([eps] * len(probs))
    return [[a.clone() for a in accs] for accs in out]


def _column_accs_sequential(pol, grid, coeffs, sources, monitor_sets, n_steps, checkpoint_every, decay):
    init = init_tm_state if pol == "tm" else init_te_state
    accs = []
    for src, mons in zip(sources, monitor_sets):
        if decay is not None:
            (ix, iy), ratio, n_decay = decay
            name = "Ez" if pol == "tm" else "Hz"
            probe = lambda st: getattr(st, name)[ix, iy] * getattr(st, name)[ix, iy]
            _, a = run_until_decayed(step_tm if pol == "tm" else step_te, init(grid), coeffs, src, mons, probe,
                                     decay_ratio=ratio, n_decay=n_decay, max_steps=n_steps)
        else:
            run = run_tm if pol == "tm" else run_te
            _, a = run(init(grid), coeffs, n_steps, source=src, monitors=mons, checkpoint_every=checkpoint_every)
        accs.append(a)
    return accs


def compute_smatrix(eps, grid, ports, *, pol, fcen, fwidth, freqs, n_steps, npml, checkpoint_every=0,
                    decay_probe=None, decay_ratio=1e-6, n_decay=50):
    """eps: cell-centred (nx, ny) map; a Gaussian pulse (fcen, fwidth) drives
    every column for n_steps steps, or, with decay_probe = (ix, iy), until the
    squared field (Ez for TM, Hz for TE) at that cell stays below
    decay_ratio times its peak for n_decay consecutive steps, n_steps at
    most.  checkpoint_every only regroups the spectra sums of sequential
    runs (chunk partials added in order)."""
    validate_ports(ports, grid, npml)
    if pol not in ("tm", "te"):
        raise ValueError(f"pol must be 'tm' or 'te', got {pol!r}")
    if pol == "te" and any(p.axis != "x" for p in ports):
        raise NotImplementedError("TE y-normal not supported (FluxMonitorTE is x-normal only)")
    coeffs = (build_tm_coeffs if pol == "tm" else build_te_coeffs)(grid, eps_inf=eps, npml=npml)
    source_cls = ModeSourceTM if pol == "tm" else ModeSourceTE
    columns = [(j, n) for j, p in enumerate(ports) for n in range(p.n_modes)]
    sources = [source_cls(grid, coeffs, ports[j], GaussianPulse(fcen, fwidth), fcen, mode=n) for j, n in columns]
    monitor_sets = [[_port_line_monitor(grid, freqs, pol, p) for p in ports] for _ in columns]
    batched = pol == "tm" and decay_probe is None and checkpoint_every <= 0 and eps.is_cuda
    if batched:
        accs = _column_accs_batched(coeffs, eps, sources, monitor_sets, n_steps)
    else:
        decay = (decay_probe, decay_ratio, n_decay) if decay_probe is not None else None
        accs = _column_accs_sequential(pol, grid, coeffs, sources, monitor_sets, n_steps, checkpoint_every, decay)

    lines = [_port_eps_line(eps, pol, p) for p in ports]
    nf, P = len(freqs), len(columns)
    s = torch.zeros(nf, P, P, dtype=torch.complex128)
    for c, (j, n) in enumerate(columns):
        # (forward, backward) amplitudes at every port for this column
        waves = [decompose(acc, line, grid.dx, freqs, pol, p.n_modes, axis_sign=_flux_sign(pol, p))
                 for p, acc, line in zip(ports, accs[c], lines)]
        launched = waves[j][0 if ports[j].direction > 0 else 1][:, n]
        received = torch.cat([waves[i][1 if p.direction > 0 else 0] for i, p in enumerate(ports)], dim=1)
        s[:, :, c] = (received / launched[:, None]).cpu()
    return SMatrix(s=s, freqs=torch.tensor([float(f) for f in freqs], dtype=torch.float64), index=columns)
