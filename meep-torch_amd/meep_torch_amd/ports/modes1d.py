"""1D slab eigenmodes on a port line (meep_torch/modes/fd1d.py:1-146).

Cell-centred nodes y_i = (i + 1/2) dx, Dirichlet (PEC) ghosts at both ends,
or a Neumann (PMC) ghost at the low end for TM half-domain ports.  With
k0 = 2 pi f and s = 1 / (k0 dx)^2:

* TM (Ez): A = D2 / k0^2 + diag(eps), symmetric tridiagonal (diagonal
  eps_i - 2 s, off-diagonal s); eigenvalue neff^2.
* TE (Hz): A = diag(eps) (L / k0^2 + I) with L = d/dy (1/eps_edge) d/dy,
  eps_edge the two-point average inside and the end cell's eps at the two
  boundary edges.  L is symmetric tridiagonal, so A v = lambda v is solved
  as the symmetric-definite pencil (L / k0^2 + I) v = lambda diag(1/eps) v,
  whose eigenvalues are real, instead of as a nonsymmetric problem; the
  eigenvectors are A's.

The k largest eigenvalues give neff = sqrt(max(lambda, 0)), descending; each
profile is scaled to unit L2 norm with its largest-magnitude entry
positive.  The solve runs on the host (scipy); the profiles come back on the
device of the eps line given."""
from typing import NamedTuple

import numpy as np
import scipy.linalg
import torch


class Modes1D(NamedTuple):
    neff: torch.Tensor       # (k,) float64, descending
    profiles: torch.Tensor   # (k, n) float64, unit L2 norm, largest entry positive


def _edge_eps(eps):
    e = np.empty(eps.size + 1)
    e[1:-1] = 0.5 * (eps[:-1] + eps[1:])
    e[0], e[-1] = eps[0], eps[-1]
    return e


def solve_1d(eps_line, dx: float, f: float, pol: str, k: int = 1, bc: str = "pec") -> Modes1D:
    if pol not in ("tm", "te"):
        raise ValueError(f"pol must be 'tm' or 'te', got {pol!r}")
    if bc not in ("pec", "pmc_lo"):
        raise ValueError(f"bc must be 'pec' or 'pmc_lo', got {bc!r}")
    if bc == "pmc_lo" and pol != "tm":
        raise ValueError("bc='pmc_lo' is implemented for pol='tm' only")
    eps = (eps_line.detach().to(torch.float64).cpu().numpy() if torch.is_tensor(eps_line)
           else np.asarray(eps_line, dtype=np.float64))
    n = eps.size
    k = min(int(k), n)
    s = 1.0 / (2.0 * np.pi * f * dx) ** 2
    sel = (n - k, n - 1)
    if pol == "tm":
        diag = eps - 2.0 * s
        if bc == "pmc_lo":
            diag[0] += s
        vals, vecs = scipy.linalg.eigh_tridiagonal(diag, np.full(n - 1, s), select="i", select_range=sel)
    else:
        w = 1.0 / _edge_eps(eps)
        B = np.diag(1.0 - s * (w[:-1] + w[1:])) + np.diag(s * w[1:-1], 1) + np.diag(s * w[1:-1], -1)
        vals, vecs = scipy.linalg.eigh(B, np.diag(1.0 / eps), subset_by_index=list(sel))
    order = np.argsort(vals)[::-1]
    vals, vecs = vals[order], vecs[:, order]
    vecs = vecs / np.linalg.norm(vecs, axis=0, keepdims=True)
    lead = vecs[np.argmax(np.abs(vecs), axis=0), np.arange(vecs.shape[1])]
    vecs = vecs * np.where(lead < 0, -1.0, 1.0)
    dev = eps_line.device if torch.is_tensor(eps_line) else torch.device("cpu")
    return Modes1D(neff=torch.from_numpy(np.sqrt(np.clip(vals, 0.0, None))).to(torch.float64),
                   profiles=torch.from_numpy(np.ascontiguousarray(vecs.T)).to(device=dev, dtype=torch.float64))
