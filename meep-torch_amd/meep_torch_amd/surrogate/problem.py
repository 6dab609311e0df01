"""The inverse-design problem the surrogate learns: broadband transmission
through a 128 x 128 design region in 2D TM, solved by the batched HIP
engine (BatchedTM: many variants per persistent launch).

Geometry (defaults; cells of dx = 0.05):
  n x n grid (n = 256), 24-cell CPML, background eps 2.1; an input and an
  output waveguide of eps 12.1 (wg cells wide) along x through the centre
  row, cut by the centred design region; eps in the region =
  2.1 + 10 * rho, rho = tanh_projection(conic_filter(theta)) of the design
  variables theta in [0, 1] (the reference's design chain,
  meep_torch/adjoint/filters.py:16-52, meep_torch/adjoint/problem.py:19-59).
  A Gaussian line source (fcen 0.6, fwidth 0.6) spans the input guide
  with a margin of one guide width each side at x = npml + 8; an x-normal
  flux line over the same span sits at x = n - npml - 8 and records 50
  frequencies in [fcen - band/2, fcen + band/2], band = 0.3 (inside the
  pulse's main lobe: source amplitude >= 0.29 of its peak there, so the
  transmission ratio stays well conditioned).
Transmission T(f) = flux(f) / flux_straight(f), the flux of the same run
with the design region filled by the straight guide.  FOM = mean_f T(f).
"""
import math

import torch

from ..adjoint.filters import conic_filter, tanh_projection
from ..batch import BatchedTM
from ..core.coeffs import build_tm_coeffs
from ..core.grid import Grid
from ..core.monitors import FluxMonitorTM
from ..core.sources import GaussianPulse, LineSource


class TransmissionProblem:
    def __init__(self, n=256, design=128, npml=24, wg=10, steps=2500, nfreq=50, fcen=0.6, fwidth=0.6,
                 band=0.3, dx=0.05, radius=0.2, beta=8.0, eps_lo=2.1, eps_hi=12.1, checkpoint_every=250,
                 device="cuda"):
        if design + 2 * (npml + 12) > n:
            raise ValueError("design region does not fit between the CPML shells")
        self.n, self.d, self.npml, self.wg, self.steps = n, design, npml, wg, steps
        self.dx, self.radius, self.beta = dx, radius, beta
        self.eps_lo, self.eps_hi = eps_lo, eps_hi
        self.K = checkpoint_every
        self.device = torch.device(device)
        self.o = (n - design) // 2
        self.j0 = n // 2 - wg // 2
        self.j1 = self.j0 + wg
        self.freqs = [fcen - band / 2 + band * i / (nfreq - 1) for i in range(nfreq)]
        self.grid = Grid(nx=n, ny=n, dx=dx, dtype=torch.float32, device=str(self.device))
        base = torch.full((n, n), eps_lo, dtype=torch.float32, device=self.device)
        base[: self.o, self.j0:self.j1] = eps_hi
        base[self.o + design:, self.j0:self.j1] = eps_hi
        self.base = base
        self.coeffs = build_tm_coeffs(self.grid, eps_inf=base, npml=npml)
        lo, hi = self.j0 - wg, self.j1 + wg
        self.source = LineSource(self.grid, ix=npml + 8, j0=lo, j1=hi,
                                 profile=GaussianPulse(fcen, fwidth, cutoff=3.0))
        self.monitor = FluxMonitorTM(self.grid, self.freqs, x_normal=True, idx=n - npml - 8, lo=lo, hi=hi)
        self._runners = {}
        self._ref = None

    # ---- design chain -------------------------------------------------------
    def project(self, theta):
        """theta (..., d, d) in [0, 1] -> projected density rho (differentiable)."""
        if theta.dim() == 2:
            return tanh_projection(conic_filter(theta, self.radius, self.dx), self.beta, 0.5)
        return torch.stack([self.project(t) for t in theta])

    def eps_of(self, rho):
        e = self.base.clone()
        o, d = self.o, self.d
        e[o:o + d, o:o + d] = self.eps_lo + (self.eps_hi - self.eps_lo) * rho
        return e

    def straight(self):
        """Density of the straight guide through the design region."""
        r = torch.zeros(self.d, self.d, dtype=torch.float32, device=self.device)
        r[:, self.j0 - self.o:self.j1 - self.o] = 1.0
        return r

    # ---- FDTD ---------------------------------------------------------------
    def runner(self, n_var, adjoint):
        key = (n_var, adjoint)
        if key not in self._runners:
            probs = [(self.coeffs, self.source, [self.monitor])] * n_var
            self._runners[key] = BatchedTM(probs, self.steps, checkpoint_every=self.K, adjoint=adjoint)
        return self._runners[key]

    def reference_flux(self):
        if self._ref is None:
            r = self.runner(1, False)
            (acc,), = r.forward([self.eps_of(self.straight())])
            self._ref = self.monitor.flux(acc, self.dx).detach().clone()
        return self._ref

    def transmission(self, rho_batch, batch=64, log=None):
        """T (count, nfreq) for a stack of projected densities (no gradient)."""
        ref = self.reference_flux()
        out = []
        with torch.no_grad():
            for s in range(0, rho_batch.shape[0], batch):
                if log and s and s % (batch * 64) == 0:
                    log(f"[surrogate] {s} / {rho_batch.shape[0]} solves")
                chunk = rho_batch[s:s + batch]
                r = self.runner(batch, False)
                eps = [self.eps_of(x) for x in chunk]
                eps += [self.base] * (batch - len(eps))   # pad the last batch
                res = r.forward(eps)
                for v in range(chunk.shape[0]):
                    out.append(self.monitor.flux(res[v][0], self.dx) / ref)
        return torch.stack(out)

    def solve_with_gradient(self, theta):
        """FDTD forward + adjoint at one design: (T (nfreq), FOM, dFOM/dtheta)."""
        ref = self.reference_flux()
        th = theta.detach().clone().requires_grad_(True)
        rho = self.project(th)
        r = self.runner(1, True)
        (acc,), = r.apply([self.eps_of(rho)])
        T = self.monitor.flux(acc, self.dx) / ref
        fom = T.mean()
        fom.backward()
        return T.detach(), float(fom.detach()), th.grad.detach()


def sample_designs(count, seed, d=128, device="cuda"):
    """Seeded design variables theta (count, d, d) in [0, 1] from three
    families: smooth random fields (correlation length 2-16 cells), a
    horizontal channel of random width/offset/bend plus noise, and mixtures
    of the two.  The chain (filter + projection) turns them into the
    densities the network sees."""
    g = torch.Generator(device=device).manual_seed(int(seed))
    u = lambda *s: torch.rand(*s, generator=g, device=device)
    out = torch.empty(count, d, d, device=device)
    ky = torch.fft.fftfreq(d, device=device).view(d, 1)
    kx = torch.fft.rfftfreq(d, device=device).view(1, -1)
    k2 = ky ** 2 + kx ** 2
    yy = torch.arange(d, device=device, dtype=torch.float32).view(1, d)
    xx = torch.arange(d, device=device, dtype=torch.float32).view(d, 1)
    for s in range(0, count, 256):
        m = min(256, count - s)
        sig = 2.0 + 14.0 * u(m, 1, 1)
        spec = torch.fft.rfft2(torch.randn(m, d, d, generator=g, device=device))
        f = torch.fft.irfft2(spec * torch.exp(-2 * (math.pi ** 2) * k2 * sig ** 2), s=(d, d))
        f = f / f.flatten(1).std(1).view(m, 1, 1).clamp_min(1e-6)
        noise = (0.5 + (0.15 + 0.35 * u(m, 1, 1)) * f + 0.3 * (u(m, 1, 1) - 0.5)).clamp(0, 1)
        width = 4.0 + 24.0 * u(m, 1, 1)
        centre = d / 2 + (u(m, 1, 1) - 0.5) * 20.0 + (u(m, 1, 1) - 0.5) * 16.0 * torch.sin(
            math.pi * xx / d * (1 + 2 * u(m, 1, 1)))
        chan = torch.sigmoid((width / 2 - (yy - centre).abs()) / 1.5)
        chan = (chan + 0.25 * (noise - 0.5)).clamp(0, 1)
        fam = u(m, 1, 1)
        w = torch.where(fam < 0.4, torch.zeros_like(fam), torch.where(fam < 0.7, torch.ones_like(fam), u(m, 1, 1)))
        out[s:s + m] = (1 - w) * noise + w * chan
    return out
