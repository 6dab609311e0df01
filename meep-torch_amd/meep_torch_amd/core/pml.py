"""CPML profiles (meep_torch/core/pml.py:8-80).

Computed once on the host CPU in the grid dtype with the reference's operation
order, so the FP32 b/a values the kernels read are bit-identical to the
reference's; then moved to the device by the coefficient builders."""
import math

import torch


def side_counts(n_or_pair) -> tuple[int, int]:
    """int -> same layer on both ends; (lo, hi) -> per end (pml.py:8-14)."""
    if isinstance(n_or_pair, (tuple, list)):
        lo, hi = n_or_pair
        return int(lo), int(hi)
    return int(n_or_pair), int(n_or_pair)


def cpml_profiles(n, npml, offset, dt, dx, m=3, R0=1e-6, alpha_max=0.05,
                  dtype=torch.float64, device="cpu"):
    """(b, a, kappa) of length n at x = i + offset (pml.py:17-59): polynomial
    sigma of order m reaching -(m+1) ln(R0) / (2 n dx) at the wall, CFS alpha
    = alpha_max (1 - u), kappa = 1, b = exp(-(sigma/kappa + alpha) dt),
    a = sigma (b - 1) / (kappa (sigma + kappa alpha)).  The high wall sits at
    n - 1 + 2 offset so E and H ramps line up."""
    n_lo, n_hi = side_counts(npml)
    x = torch.arange(n, dtype=dtype) + offset
    u = torch.zeros(n, dtype=dtype)
    sig = torch.zeros_like(u)
    if n_lo > 0:
        u_lo = torch.clamp((n_lo - x) / n_lo, min=0.0)
        u = u + u_lo
        sig = sig + (-(m + 1) * math.log(R0) / (2.0 * n_lo * dx)) * u_lo ** m
    if n_hi > 0:
        wall = n - 1 + 2.0 * offset
        u_hi = torch.clamp((x - (wall - n_hi)) / n_hi, min=0.0)
        u = u + u_hi
        sig = sig + (-(m + 1) * math.log(R0) / (2.0 * n_hi * dx)) * u_hi ** m
    u = torch.clamp(u, max=1.0)
    alpha = torch.where(u > 0, alpha_max * (1.0 - u), torch.zeros_like(u))
    kappa = torch.ones_like(u)
    b = torch.exp(-(sig / kappa + alpha) * dt)
    denom = kappa * (sig + kappa * alpha)
    a = torch.where(sig > 0, sig * (b - 1.0) / denom.clamp(min=1e-300), torch.zeros_like(b))
    return b.to(device), a.to(device), kappa.to(device)


def absorber_sigma(n, nabs, dx, *, R0=1e-15, m=2, dtype=torch.float64, device="cpu"):
    """Matched-absorber sigma ramp (pml.py:62-80): on each absorbing edge of
    nc cells, smax ((nc - i - 1/2) / nc)^m from the wall inwards, with
    smax = -(m + 1) ln(R0) / (2 nc dx); zero elsewhere.  Feeds the 3D
    absorber boundary (build_3d_coeffs n_abs) and the facade's Absorber
    layers."""
    n_lo, n_hi = side_counts(nabs)
    p = torch.zeros(n, dtype=dtype)

    def ramp(nc):
        smax = -(m + 1) * math.log(R0) / (2.0 * nc * dx)
        i = torch.arange(nc, dtype=dtype)
        return smax * ((nc - i - 0.5) / nc) ** m

    if n_lo > 0:
        p[:n_lo] = ramp(n_lo)
    if n_hi > 0:
        p[n - n_hi:] = torch.flip(ramp(n_hi), [0])
    return p.to(device)
