"""Density filters and projection (meep_torch/adjoint/filters.py:16-70).

Edge-renormalised correlation (constant fields are invariant), conic and
Gaussian kernels, tanh threshold projection.  Small conv2d/conv3d work on the
design grid, left to PyTorch (MIOpen on the device); the FDTD step is the hot
path, not these."""
import math

import torch
import torch.nn.functional as F


def _conv_norm(rho, kern):
    k = kern[None, None]
    x = rho[None, None]
    pad = (kern.shape[0] // 2, kern.shape[1] // 2)
    num = F.conv2d(x, k, padding=pad)
    den = F.conv2d(torch.ones_like(x), k, padding=pad)
    return (num / den)[0, 0]


def conic_filter(rho, radius: float, dx: float):
    """Linear cone of radius `radius` (physical units), filters.py:27-33."""
    n = max(1, int(math.ceil(radius / dx)))
    ax = torch.arange(-n, n + 1, dtype=rho.dtype, device=rho.device) * dx
    X, Y = torch.meshgrid(ax, ax, indexing="ij")
    kern = torch.clamp(1.0 - torch.sqrt(X ** 2 + Y ** 2) / radius, min=0.0)
    return _conv_norm(rho, kern)


def gaussian_filter(rho, sigma: float, dx: float):
    """Gaussian of standard deviation `sigma` (physical units), filters.py:36-42."""
    n = max(1, int(math.ceil(3 * sigma / dx)))
    ax = torch.arange(-n, n + 1, dtype=rho.dtype, device=rho.device) * dx
    X, Y = torch.meshgrid(ax, ax, indexing="ij")
    kern = torch.exp(-(X ** 2 + Y ** 2) / (2 * sigma ** 2))
    return _conv_norm(rho, kern)


def tanh_projection(rho, beta: float, eta: float = 0.5):
    """Smoothed Heaviside at eta with sharpness beta, filters.py:45-52."""
    num = torch.tanh(torch.as_tensor(beta * eta, dtype=rho.dtype, device=rho.device)) \
        + torch.tanh(beta * (rho - eta))
    den = math.tanh(beta * eta) + math.tanh(beta * (1 - eta))
    return num / den


def _conv_norm3d(rho, kern):
    k = kern[None, None]
    x = rho[None, None]
    pad = tuple(s // 2 for s in kern.shape)
    num = F.conv3d(x, k, padding=pad)
    den = F.conv3d(torch.ones_like(x), k, padding=pad)
    return (num / den)[0, 0]


def conic_filter3d(rho, radius: float, dx: float):
    """3D cone blur, filters.py:64-70."""
    n = max(1, int(math.ceil(radius / dx)))
    ax = torch.arange(-n, n + 1, dtype=rho.dtype, device=rho.device) * dx
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    kern = torch.clamp(1.0 - torch.sqrt(X ** 2 + Y ** 2 + Z ** 2) / radius, min=0.0)
    return _conv_norm3d(rho, kern)
