"""TM/TE coefficient tuples (meep_torch/core/coeffs.py:21-74, 294-345) and
their builders (coeffs.py:134-291, 357-503).

The tuples keep the reference's field names and shapes so code that reads
them keeps working; one extra trailing field, ``shells``, carries the CPML
shell bounds the kernels use to skip the psi recursion in the interior.

Supported by the HIP stepper: eps_inf, npml (int or ((xlo, xhi), (ylo, yhi))),
and y_pmc for TM.  Dispersive poles, chi2/chi3, sigma_d/sigma_b/mu, magnetic
poles, tensor eps and symmetry folds raise NotImplementedError naming the
flag (no silent fallback)."""
from typing import NamedTuple, Sequence

import torch

from .grid import Grid
from .pml import cpml_profiles, side_counts
from .._engine import shell_bounds


def _axis_sides(npml):
    if isinstance(npml, (tuple, list)):
        px, py = npml
        return side_counts(px), side_counts(py)
    return (int(npml), int(npml)), (int(npml), int(npml))


class TMCoeffs(NamedTuple):
    dt: float
    dx: float
    eps_inf: torch.Tensor
    chi3: torch.Tensor
    has_kerr: bool
    c1: torch.Tensor
    c2: torch.Tensor
    c3: torch.Tensor
    y_pmc: bool
    b_hy: torch.Tensor
    a_hy: torch.Tensor
    k_hy: torch.Tensor
    b_hx: torch.Tensor
    a_hx: torch.Tensor
    k_hx: torch.Tensor
    b_ex: torch.Tensor
    a_ex: torch.Tensor
    k_ex: torch.Tensor
    b_ey: torch.Tensor
    a_ey: torch.Tensor
    k_ey: torch.Tensor
    d1: torch.Tensor | None = None
    d2: torch.Tensor | None = None
    h1x: torch.Tensor | None = None
    h2x: torch.Tensor | None = None
    h1y: torch.Tensor | None = None
    h2y: torch.Tensor | None = None
    mc1x: torch.Tensor | None = None
    mc2x: torch.Tensor | None = None
    mc3x: torch.Tensor | None = None
    mc1y: torch.Tensor | None = None
    mc2y: torch.Tensor | None = None
    mc3y: torch.Tensor | None = None
    chi2: torch.Tensor | None = None
    has_chi2: bool = False
    y_mirror_lo: int = 0
    x_mirror_lo: int = 0
    y_rot2_lo: int = 0
    rot4_lo: int = 0
    shells: dict | None = None


class TECoeffs(NamedTuple):
    dt: float
    dx: float
    eps_x: torch.Tensor
    eps_y: torch.Tensor
    chi3_x: torch.Tensor
    chi3_y: torch.Tensor
    has_kerr: bool
    c1x: torch.Tensor
    c2x: torch.Tensor
    c3x: torch.Tensor
    c1y: torch.Tensor
    c2y: torch.Tensor
    c3y: torch.Tensor
    b_hx: torch.Tensor
    a_hx: torch.Tensor
    k_hx: torch.Tensor
    b_hy: torch.Tensor
    a_hy: torch.Tensor
    k_hy: torch.Tensor
    b_ey: torch.Tensor
    a_ey: torch.Tensor
    k_ey: torch.Tensor
    b_ex: torch.Tensor
    a_ex: torch.Tensor
    k_ex: torch.Tensor
    d1x: torch.Tensor | None = None
    d2x: torch.Tensor | None = None
    d1y: torch.Tensor | None = None
    d2y: torch.Tensor | None = None
    hz1: torch.Tensor | None = None
    hz2: torch.Tensor | None = None
    inv_xx_x: torch.Tensor | None = None
    inv_xy_x: torch.Tensor | None = None
    inv_yy_y: torch.Tensor | None = None
    inv_xy_y: torch.Tensor | None = None
    chi2_x: torch.Tensor | None = None
    chi2_y: torch.Tensor | None = None
    has_chi2: bool = False
    c3xy: torch.Tensor | None = None
    c3yx: torch.Tensor | None = None
    y_mirror_lo: int = 0
    x_mirror_lo: int = 0
    shells: dict | None = None


def _as_field(val, grid: Grid, default: float) -> torch.Tensor:
    """coeffs.py:77-84: scalar -> filled grid, tensor -> shape-checked cast."""
    if val is None:
        val = default
    if isinstance(val, torch.Tensor):
        if val.shape != (grid.nx, grid.ny):
            raise ValueError(f"expected shape {(grid.nx, grid.ny)}, got {tuple(val.shape)}")
        return val.to(dtype=grid.dtype, device=grid.device)
    return torch.full((grid.nx, grid.ny), float(val), dtype=grid.dtype, device=grid.device)


def _unsupported(flag):
    raise NotImplementedError(
        f"meep_torch_amd HIP stepper: '{flag}' is not supported yet (eps, CPML and y_pmc are)")


def _profile(grid, n, offset, pair):
    """Profile on the host in the grid dtype (bit-identical to the reference),
    shell bounds, then the device copy."""
    b, a, k = cpml_profiles(n, pair, offset, grid.dt, grid.dx, dtype=grid.dtype, device="cpu")
    return b, a, k


def build_tm_coeffs(grid: Grid, eps_inf=None, poles: Sequence = (), pole_masks=None,
                    chi3=None, chi2=None, npml=0, y_pmc: bool = False, sigma_d=None,
                    sigma_b=None, mu=None, m_poles: Sequence = (), m_pole_masks=None,
                    y_mirror: int = 0, x_mirror: int = 0, y_rot2: int = 0,
                    rot4: int = 0) -> TMCoeffs:
    """coeffs.py:134-291, eps + CPML (+ y_pmc) subset."""
    dt, dx = grid.dt, grid.dx
    eps = _as_field(eps_inf, grid, 1.0)
    x3 = _as_field(chi3, grid, 0.0)
    x2 = _as_field(chi2, grid, 0.0)
    if bool((x3 != 0).any()):
        _unsupported("chi3 (Kerr)")
    if bool((x2 != 0).any()):
        _unsupported("chi2")
    if len(poles):
        _unsupported("poles (ADE dispersion)")
    if sigma_d is not None or sigma_b is not None or mu is not None:
        _unsupported("sigma_d/sigma_b/mu")
    if len(m_poles):
        _unsupported("m_poles")
    if y_mirror or x_mirror or y_rot2 or rot4:
        _unsupported("symmetry folds (y_mirror/x_mirror/y_rot2/rot4)")
    px, py = _axis_sides(npml)
    if sum(px) >= grid.nx or (sum(py) >= grid.ny and not y_pmc):
        raise ValueError(f"npml={npml} too large for grid {grid.nx}x{grid.ny}")

    b_hy, a_hy, k_hy = _profile(grid, grid.ny - 1, 0.5, (0, 0) if y_pmc else py)
    b_hx, a_hx, k_hx = _profile(grid, grid.nx - 1, 0.5, px)
    b_exf, a_exf, k_exf = _profile(grid, grid.nx, 0.0, px)
    b_ex, a_ex, k_ex = b_exf[1:-1], a_exf[1:-1], k_exf[1:-1]
    b_eyf, a_eyf, k_eyf = _profile(grid, grid.ny, 0.0, (0, 0) if y_pmc else py)
    if y_pmc:
        b_ey, a_ey, k_ey = b_eyf, a_eyf, k_eyf
    else:
        b_ey, a_ey, k_ey = b_eyf[1:-1], a_eyf[1:-1], k_eyf[1:-1]
    shells = {"hy": shell_bounds(b_hy, a_hy), "hx": shell_bounds(b_hx, a_hx),
              "ex": shell_bounds(b_ex, a_ex), "ey": shell_bounds(b_ey, a_ey)}
    dev = lambda t: t.contiguous().to(grid.device)
    z = grid.zeros(0, grid.nx, grid.ny)
    return TMCoeffs(
        dt=dt, dx=dx, eps_inf=eps, chi3=x3, has_kerr=False, chi2=x2, has_chi2=False,
        c1=z, c2=z.clone(), c3=z.clone(), y_pmc=y_pmc,
        b_hy=dev(b_hy)[None, :], a_hy=dev(a_hy)[None, :], k_hy=dev(k_hy)[None, :],
        b_hx=dev(b_hx)[:, None], a_hx=dev(a_hx)[:, None], k_hx=dev(k_hx)[:, None],
        b_ex=dev(b_ex)[:, None], a_ex=dev(a_ex)[:, None], k_ex=dev(k_ex)[:, None],
        b_ey=dev(b_ey)[None, :], a_ey=dev(a_ey)[None, :], k_ey=dev(k_ey)[None, :],
        shells=shells)


def _avg_to_ex(v):
    return 0.5 * (v[:-1, :] + v[1:, :])


def _avg_to_ey(v):
    return 0.5 * (v[:, :-1] + v[:, 1:])


def build_te_coeffs(grid: Grid, eps_inf=None, poles: Sequence = (), pole_masks=None,
                    chi3=None, chi2=None, npml=0, sigma_d=None, sigma_b=None, mu=None,
                    eps_tensor=None, poles_x: Sequence = (), poles_y: Sequence = (),
                    pole_masks_x=None, pole_masks_y=None, poles_tensor: Sequence = (),
                    pole_masks_tensor=None, y_mirror: int = 0, x_mirror: int = 0) -> TECoeffs:
    """coeffs.py:357-503, eps + CPML subset.  eps_x / eps_y are the two-point
    averages of the cell eps (coeffs.py:347-355,385)."""
    dt, dx = grid.dt, grid.dx
    eps = _as_field(eps_inf, grid, 1.0)
    x3 = _as_field(chi3, grid, 0.0)
    x2 = _as_field(chi2, grid, 0.0)
    if bool((x3 != 0).any()):
        _unsupported("chi3 (Kerr)")
    if bool((x2 != 0).any()):
        _unsupported("chi2")
    if len(poles) or len(poles_x) or len(poles_y) or len(poles_tensor):
        _unsupported("poles (ADE dispersion)")
    if sigma_d is not None or sigma_b is not None or mu is not None:
        _unsupported("sigma_d/sigma_b/mu")
    if eps_tensor is not None:
        _unsupported("eps_tensor (anisotropic eps)")
    if y_mirror or x_mirror:
        _unsupported("symmetry folds (y_mirror/x_mirror)")
    eps_x, eps_y = _avg_to_ex(eps), _avg_to_ey(eps)
    px, py = _axis_sides(npml)
    if sum(px) >= grid.nx or sum(py) >= grid.ny:
        raise ValueError(f"npml={npml} too large for grid {grid.nx}x{grid.ny}")
    b_hx, a_hx, k_hx = _profile(grid, grid.nx - 1, 0.5, px)
    b_hy, a_hy, k_hy = _profile(grid, grid.ny - 1, 0.5, py)
    b_eyf, a_eyf, k_eyf = _profile(grid, grid.ny, 0.0, py)
    b_exf, a_exf, k_exf = _profile(grid, grid.nx, 0.0, px)
    b_ey, a_ey, k_ey = b_eyf[1:-1], a_eyf[1:-1], k_eyf[1:-1]
    b_ex, a_ex, k_ex = b_exf[1:-1], a_exf[1:-1], k_exf[1:-1]
    shells = {"hx": shell_bounds(b_hx, a_hx), "hy": shell_bounds(b_hy, a_hy),
              "ey": shell_bounds(b_ey, a_ey), "ex": shell_bounds(b_ex, a_ex)}
    dev = lambda t: t.contiguous().to(grid.device)
    zx = grid.zeros(0, grid.nx - 1, grid.ny)
    zy = grid.zeros(0, grid.nx, grid.ny - 1)
    return TECoeffs(
        dt=dt, dx=dx, eps_x=eps_x, eps_y=eps_y, chi3_x=_avg_to_ex(x3), chi3_y=_avg_to_ey(x3),
        has_kerr=False, chi2_x=_avg_to_ex(x2), chi2_y=_avg_to_ey(x2), has_chi2=False,
        c1x=zx, c2x=zx.clone(), c3x=zx.clone(), c1y=zy, c2y=zy.clone(), c3y=zy.clone(),
        b_hx=dev(b_hx)[:, None], a_hx=dev(a_hx)[:, None], k_hx=dev(k_hx)[:, None],
        b_hy=dev(b_hy)[None, :], a_hy=dev(a_hy)[None, :], k_hy=dev(k_hy)[None, :],
        b_ey=dev(b_ey)[None, :], a_ey=dev(a_ey)[None, :], k_ey=dev(k_ey)[None, :],
        b_ex=dev(b_ex)[:, None], a_ex=dev(a_ex)[:, None], k_ex=dev(k_ex)[:, None],
        shells=shells)
