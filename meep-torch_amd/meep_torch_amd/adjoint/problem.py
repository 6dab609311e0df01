"""Density-based design chain in the meep.adjoint shape
(meep_torch/adjoint/problem.py:19-110):

    rho in [0, 1] -> conic filter -> tanh projection -> eps_min + (eps_max -
    eps_min) * rho_p, bilinearly upsampled into an index box of the base eps
    -> user run function -> scalar objective J.

`OptimizationProblem()` returns (J, dJ/drho).  When the run function drives
`meep_torch_amd.core.run.run_tm`, the FDTD part of the gradient comes from
the hand-written reverse-time adjoint kernels (the eps leaf requires grad),
and torch autograd only carries the cheap filter/projection/paint chain.
"""
import torch
import torch.nn.functional as F

from .filters import conic_filter, conic_filter3d, tanh_projection


class MaterialGrid:
    """Design grid of `shape`; rho linearly interpolates eps_min..eps_max."""

    def __init__(self, shape, eps_min=1.0, eps_max=12.0, beta=0.0, eta=0.5):
        self.shape = tuple(shape)
        self.eps_min, self.eps_max = float(eps_min), float(eps_max)
        self.beta, self.eta = beta, eta

    def eps(self, rho):
        return self.eps_min + (self.eps_max - self.eps_min) * rho


class DesignRegion:
    """Paints a MaterialGrid into eps[ix0:ix1, iy0:iy1] (bilinear,
    align_corners=False, as torch.nn.functional.interpolate defines it)."""

    def __init__(self, material_grid, *, ix0, ix1, iy0, iy1):
        if ix1 <= ix0 or iy1 <= iy0:
            raise ValueError("empty design region")
        self.mg = material_grid
        self.box = (ix0, ix1, iy0, iy1)

    def paint(self, base_eps, rho_projected):
        ix0, ix1, iy0, iy1 = self.box
        up = F.interpolate(rho_projected[None, None], size=(ix1 - ix0, iy1 - iy0), mode="bilinear",
                           align_corners=False)[0, 0]
        out = base_eps.clone()
        out[ix0:ix1, iy0:iy1] = self.mg.eps(up)
        return out


class DesignRegion3D:
    def __init__(self, material_grid, *, ix0, ix1, iy0, iy1, iz0, iz1):
        self.mg = material_grid
        self.box = (ix0, ix1, iy0, iy1, iz0, iz1)

    def paint(self, base_eps, rho_projected):
        ix0, ix1, iy0, iy1, iz0, iz1 = self.box
        up = F.interpolate(rho_projected[None, None], size=(ix1 - ix0, iy1 - iy0, iz1 - iz0),
                           mode="trilinear", align_corners=False)[0, 0]
        out = base_eps.clone()
        out[ix0:ix1, iy0:iy1, iz0:iz1] = self.mg.eps(up)
        return out


def rho_from_eps(eps, region):
    """Seed density from an eps map: the region box normalised by the
    MaterialGrid endpoints, area-downsampled to the design grid, clamped to
    [0.02, 0.98]."""
    ix0, ix1, iy0, iy1 = region.box
    mg = region.mg
    r = (eps[ix0:ix1, iy0:iy1] - mg.eps_min) / (mg.eps_max - mg.eps_min)
    r = F.interpolate(r[None, None].to(torch.float64), size=mg.shape, mode="area")[0, 0]
    return r.clamp(0.02, 0.98)


class OptimizationProblem:
    """opt() -> (J, dJ/drho as numpy).  `opt.rho` holds the design; update it
    between calls.  beta/eta default to the MaterialGrid's (beta 0 means no
    sharpening, taken as 1), and may be reassigned for a beta schedule."""

    def __init__(self, run_fn, base_eps, design_region, *, filter_radius=0.0, dx=1.0, beta=None, eta=None,
                 rho0=None):
        self.run_fn = run_fn
        self.base_eps = base_eps
        self.region = design_region
        self.filter_radius = float(filter_radius)
        self.dx = dx
        mg = design_region.mg
        self.beta = beta if beta is not None else (mg.beta or 1.0)
        self.eta = eta if eta is not None else (mg.eta if mg.eta is not None else 0.5)
        self.rho = rho0.clone() if rho0 is not None else torch.full(mg.shape, 0.5, dtype=base_eps.dtype,
                                                                    device=base_eps.device)

    def _filtered(self, rho):
        if self.filter_radius > 0:
            rho = conic_filter3d(rho, self.filter_radius, self.dx) if rho.dim() == 3 else \
                conic_filter(rho, self.filter_radius, self.dx)
        return rho

    def _objective(self, rho):
        r = tanh_projection(self._filtered(rho), self.beta, self.eta)
        return self.run_fn(self.region.paint(self.base_eps, r))

    def __call__(self, need_gradient=True):
        rho = self.rho.detach().clone().requires_grad_(need_gradient)
        J = self._objective(rho)
        if not need_gradient:
            return float(J.detach()), None
        J.backward()
        return float(J.detach()), rho.grad.detach().cpu().numpy().copy()

    def eval_at_delta(self, i, j, h):
        rho = self.rho.detach().clone()
        rho[i, j] += h
        with torch.no_grad():
            return float(self._objective(rho))
