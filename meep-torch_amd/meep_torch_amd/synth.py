"""Synthetic workloads of the benchmark configurations, regenerated bit-for-bit
from a counter-based generator (splitmix64) so the GPU box needs no data files.

C1  2D TM 256x256, dx=0.05, Courant 0.5, eps=12 guide rows [123,133), 16-cell
    PML, point Ez source at (24,128), GaussianPulse(0.15, 0.1), 2000 steps,
    50-frequency flux line x=232 rows [96,160).
C2  2D TE 2048x2048, dx=0.05, density U[0,1) (seed 0) on the centred 512x512,
    Gaussian filter sigma = 5 cells, eps = 2.1 + 10 rho (2.1 elsewhere), 32-cell
    CPML, Jy line source at ix=64 rows [32,2016), GaussianPulse(0.5, 0.2),
    10000 steps, flux lines x=200, x=1848 (Ey/Hz) and y=200, y=1848 (Ex/Hz),
    rows [32,2016), 100 frequencies in [0.4, 0.6].
C3  3D 384^3, dx=0.05, SiO2 eps=2.1 below z0, Si slab eps=12.1 (22 cells thick
    in z, 50 wide in y, along x), 12-cell CPML, Ez point dipole at the guide
    input, GaussianPulse(0.5, 0.2), 2000 steps, 2 x-normal flux planes x 64
    frequencies in [0.4, 0.6].
"""

import numpy as np
import torch

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix_uniform(n: int, seed: int) -> np.ndarray:
    """U[0,1) doubles: z_i = seed + (i+1)*golden, splitmix64 finaliser, top 53 bits."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def c1_eps(dtype=torch.float32):
    eps = torch.ones(256, 256, dtype=dtype)
    eps[:, 123:133] = 12.0
    return eps


def c2_eps(n=2048, region=512, seed=0, sigma_cells=5.0, dx=0.05, device="cpu",
           dtype=torch.float32):
    from .adjoint.filters import gaussian_filter
    rho = torch.tensor(splitmix_uniform(region * region, seed).reshape(region, region),
                       dtype=torch.float64, device=device)
    rho = gaussian_filter(rho, sigma_cells * dx, dx)
    eps = torch.full((n, n), 2.1, dtype=torch.float64, device=device)
    o = (n - region) // 2
    eps[o:o + region, o:o + region] = 2.1 + 10.0 * rho
    return eps.to(dtype)


C2 = dict(n=2048, dx=0.05, npml=32, steps=10000, src_ix=64, src_lo=32, src_hi=2016,
          fcen=0.5, fwidth=0.2, lines_x=(200, 1848), lines_y=(200, 1848), line_lo=32,
          line_hi=2016, nfreq=100)


def c2_freqs():
    return np.linspace(C2["fcen"] - 0.1, C2["fcen"] + 0.1, C2["nfreq"]).tolist()


def c3_eps(n=384, dtype=torch.float32):
    eps = torch.full((n, n, n), 1.0, dtype=dtype)
    z0 = n // 2 - 11
    eps[:, :, :z0] = 2.1
    eps[:, n // 2 - 25:n // 2 + 25, z0:z0 + 22] = 12.1
    return eps


C3 = dict(n=384, dx=0.05, npml=12, steps=2000, fcen=0.5, fwidth=0.2, planes=(60, 324), nfreq=64)
