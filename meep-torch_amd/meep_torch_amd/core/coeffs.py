"""TM/TE coefficient tuples (meep_torch/core/coeffs.py:21-74, 294-345) and
their builders (coeffs.py:134-291, 357-503).

The tuples keep the reference's field names and shapes so code that reads
them keeps working; one extra trailing field, ``shells``, carries the CPML
shell bounds the kernels use to skip the psi recursion in the interior.

Supported by the HIP stepper: eps_inf, npml (int or ((xlo, xhi), (ylo, yhi))),
y_pmc for TM, the media terms (electric ADE poles, chi2/chi3, sigma_d/sigma_b/
mu; TM magnetic poles; TE per-axis poles, tensor eps and oriented poles) and
the symmetry cuts (TM mirrors, Rotate2, Rotate4; TE mirrors).  Combinations
the reference refuses raise as it does."""
import math
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


def pole_recurrence(pole, dt):
    """coeffs.py:87-106: ADE scalars of one pole (fn, gamma, sigma[, fc]);
    fc = fn is a Lorentz pole, fn = 0 with fc the plasma frequency a Drude
    pole.  w0 = 2 pi fn, wc = 2 pi fc, G = 2 pi gamma, den = 1 + G dt / 2:
    c1 = (2 - w0^2 dt^2) / den, c2 = -(1 - G dt / 2) / den,
    c3 = sigma wc^2 dt^2 / den (times the pole's spatial mask)."""
    fn, gamma, sigma = pole[0], pole[1], pole[2]
    fc = pole[3] if len(pole) > 3 else fn
    w0, wc, G = 2 * math.pi * fn, 2 * math.pi * fc, 2 * math.pi * gamma
    den = 1.0 + G * dt / 2.0
    return (2.0 - w0 * w0 * dt * dt) / den, -(1.0 - G * dt / 2.0) / den, sigma * wc * wc * dt * dt / den


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
    """coeffs.py:134-291: eps + CPML (+ y_pmc or symmetry cuts at the low
    edges: y_mirror / x_mirror parity, y_rot2 and rot4 phases), electric ADE poles
    (coeffs.py:87-106, 190-198), chi2 / chi3 (coeffs.py:154-158),
    sigma_d / sigma_b / mu (coeffs.py:160-175) and magnetic Lorentz poles
    m_poles (fn, gamma, sigma) with cell masks m_pole_masks
    (coeffs.py:177-196)."""
    dt, dx = grid.dt, grid.dx
    eps = _as_field(eps_inf, grid, 1.0)
    x3 = _as_field(chi3, grid, 0.0)
    x2 = _as_field(chi2, grid, 0.0)
    has_kerr = bool((x3 != 0).any())
    has_chi2 = bool((x2 != 0).any())
    cond = {}
    if sigma_d is not None or sigma_b is not None or mu is not None:
        sd = _as_field(sigma_d, grid, 0.0)
        sb = _as_field(sigma_b, grid, 0.0)
        mu_ = _as_field(mu, grid, 1.0)
        cond["d1"] = (1 - sd * dt / 2) / (1 + sd * dt / 2)
        cond["d2"] = dt / (1 + sd * dt / 2)
        sbx, mux = 0.5 * (sb[:, :-1] + sb[:, 1:]), 0.5 * (mu_[:, :-1] + mu_[:, 1:])   # Hx (nx, ny-1)
        sby, muy = 0.5 * (sb[:-1, :] + sb[1:, :]), 0.5 * (mu_[:-1, :] + mu_[1:, :])   # Hy (nx-1, ny)
        cond["h1x"] = (1 - sbx * dt / 2) / (1 + sbx * dt / 2)
        cond["h2x"] = dt / (mux * (1 + sbx * dt / 2))
        cond["h1y"] = (1 - sby * dt / 2) / (1 + sby * dt / 2)
        cond["h2y"] = dt / (muy * (1 + sby * dt / 2))
    n_poles = len(poles)
    c1 = grid.zeros(n_poles, grid.nx, grid.ny)
    c2 = grid.zeros(n_poles, grid.nx, grid.ny)
    c3 = grid.zeros(n_poles, grid.nx, grid.ny)
    for k, pole in enumerate(poles):
        v1, v2, v3 = pole_recurrence(pole, dt)
        mask = pole_masks[k] if pole_masks is not None else torch.ones_like(eps)
        c1[k] = v1 * torch.ones_like(eps)
        c2[k] = v2 * torch.ones_like(eps)
        c3[k] = v3 * mask.to(eps)
    if len(m_poles):
        # coeffs.py:177-196: magnetic Lorentz poles on the Hx / Hy grids, the
        # per-pole scalars broadcast and mc3 scaled by the mask averaged onto
        # the H component (not combined with sigma_b / mu)
        if cond:
            raise NotImplementedError("m_poles with sigma_b/mu not combined (YAGNI)")
        nm = len(m_poles)
        for suf, shape, avg in (("x", (grid.nx, grid.ny - 1), lambda v: 0.5 * (v[:, :-1] + v[:, 1:])),
                                ("y", (grid.nx - 1, grid.ny), lambda v: 0.5 * (v[:-1, :] + v[1:, :]))):
            c1m, c2m, c3m = grid.zeros(nm, *shape), grid.zeros(nm, *shape), grid.zeros(nm, *shape)
            for k, pole in enumerate(m_poles):
                w0, G = 2 * math.pi * pole[0], 2 * math.pi * pole[1]
                den = 1.0 + G * dt / 2.0
                mask = (m_pole_masks[k] if m_pole_masks is not None
                        else torch.ones(grid.nx, grid.ny, dtype=grid.dtype, device=grid.device))
                c1m[k] = (2.0 - w0 * w0 * dt * dt) / den
                c2m[k] = -(1.0 - G * dt / 2.0) / den
                c3m[k] = (pole[2] * w0 * w0 * dt * dt / den) * avg(mask.to(device=grid.device, dtype=grid.dtype))
            cond[f"mc1{suf}"], cond[f"mc2{suf}"], cond[f"mc3{suf}"] = c1m, c2m, c3m
    # symmetry cuts (coeffs.py:211-281): a cut replaces the low-side CPML of
    # its axis, and the D-update profile keeps nodes 0..n-2 (the cut node is live)
    ycut, xcut = bool(y_mirror or y_rot2 or rot4), bool(x_mirror or rot4)
    px, py = _axis_sides(npml)
    if ycut:
        py = (0, py[1])
    if xcut:
        px = (0, px[1])
    if sum(px) >= grid.nx or (sum(py) >= grid.ny and not y_pmc):
        raise ValueError(f"npml={npml} too large for grid {grid.nx}x{grid.ny}")
    if (y_mirror not in (0, +1, -1) or x_mirror not in (0, +1, -1)
            or y_rot2 not in (0, +1, -1) or rot4 not in (0, +1, -1)):
        raise ValueError(f"symmetry parity must be 0/+1/-1, got y={y_mirror} "
                         f"x={x_mirror} rot2={y_rot2} rot4={rot4}")
    if y_mirror and y_rot2:
        raise ValueError("y_mirror and y_rot2 both cut the low-y edge")
    if rot4 and (y_mirror or x_mirror or y_rot2):
        raise ValueError("rot4 cuts both low edges on its own")
    if rot4 and grid.nx != grid.ny:
        raise ValueError(f"rot4 needs a square grid, got {grid.nx}x{grid.ny}")
    if y_pmc and (ycut or xcut):
        raise ValueError("symmetry cuts and y_pmc are mutually exclusive")
    if (y_mirror == -1 or x_mirror == -1 or y_rot2 or rot4) and ("h1x" in cond or m_poles):
        raise NotImplementedError(
            "odd mirror / Rotate2 / Rotate4 with sigma_b/mu/m_poles: the on-plane H row "
            "has no conductivity/mu coefficients (standard H-grid shapes only)")

    b_hy, a_hy, k_hy = _profile(grid, grid.ny - 1, 0.5, (0, 0) if y_pmc else py)
    b_hx, a_hx, k_hx = _profile(grid, grid.nx - 1, 0.5, px)
    b_exf, a_exf, k_exf = _profile(grid, grid.nx, 0.0, px)
    if xcut:
        b_ex, a_ex, k_ex = b_exf[:-1], a_exf[:-1], k_exf[:-1]
    else:
        b_ex, a_ex, k_ex = b_exf[1:-1], a_exf[1:-1], k_exf[1:-1]
    b_eyf, a_eyf, k_eyf = _profile(grid, grid.ny, 0.0, (0, 0) if y_pmc else py)
    if y_pmc:
        b_ey, a_ey, k_ey = b_eyf, a_eyf, k_eyf
    elif ycut:
        b_ey, a_ey, k_ey = b_eyf[:-1], a_eyf[:-1], k_eyf[:-1]
    else:
        b_ey, a_ey, k_ey = b_eyf[1:-1], a_eyf[1:-1], k_eyf[1:-1]
    shells = {"hy": shell_bounds(b_hy, a_hy), "hx": shell_bounds(b_hx, a_hx),
              "ex": shell_bounds(b_ex, a_ex), "ey": shell_bounds(b_ey, a_ey)}
    dev = lambda t: t.contiguous().to(grid.device)
    return TMCoeffs(
        dt=dt, dx=dx, eps_inf=eps, chi3=x3, has_kerr=has_kerr, chi2=x2, has_chi2=has_chi2,
        c1=c1, c2=c2, c3=c3, y_pmc=y_pmc, y_mirror_lo=y_mirror, x_mirror_lo=x_mirror, y_rot2_lo=y_rot2,
        rot4_lo=rot4, **cond,
        b_hy=dev(b_hy)[None, :], a_hy=dev(a_hy)[None, :], k_hy=dev(k_hy)[None, :],
        b_hx=dev(b_hx)[:, None], a_hx=dev(a_hx)[:, None], k_hx=dev(k_hx)[:, None],
        b_ex=dev(b_ex)[:, None], a_ex=dev(a_ex)[:, None], k_ex=dev(k_ex)[:, None],
        b_ey=dev(b_ey)[None, :], a_ey=dev(a_ey)[None, :], k_ey=dev(k_ey)[None, :],
        shells=shells)


def _axis_poles(grid, dt, eps_ref, shared, axis_poles, shared_masks, axis_masks, shape, avg):
    """coeffs.py:109-131: (c1, c2, c3) of one TE E axis; the shared poles
    come first, then this axis' own; each pole's cell mask (ones when not
    given) is Yee-averaged onto the axis by ``avg``."""
    allp = list(shared) + list(axis_poles)
    c1 = grid.zeros(len(allp), *shape)
    c2 = grid.zeros(len(allp), *shape)
    c3 = grid.zeros(len(allp), *shape)
    ns = len(shared)
    for k, pole in enumerate(allp):
        v1, v2, v3 = pole_recurrence(pole, dt)
        if k < ns:
            m = shared_masks[k] if shared_masks is not None else None
        else:
            m = axis_masks[k - ns] if axis_masks is not None else None
        mask = (m if m is not None else torch.ones_like(eps_ref)).to(eps_ref)
        c1[k] = v1
        c2[k] = v2
        c3[k] = v3 * avg(mask)
    return c1, c2, c3


def _avg_to_ex(v):
    return 0.5 * (v[:-1, :] + v[1:, :])


def _avg_to_ey(v):
    return 0.5 * (v[:, :-1] + v[:, 1:])


def build_te_coeffs(grid: Grid, eps_inf=None, poles: Sequence = (), pole_masks=None,
                    chi3=None, chi2=None, npml=0, sigma_d=None, sigma_b=None, mu=None,
                    eps_tensor=None, poles_x: Sequence = (), poles_y: Sequence = (),
                    pole_masks_x=None, pole_masks_y=None, poles_tensor: Sequence = (),
                    pole_masks_tensor=None, y_mirror: int = 0, x_mirror: int = 0) -> TECoeffs:
    """coeffs.py:357-503: eps + CPML, electric poles shared by both axes and
    per-axis (poles_x / poles_y, coeffs.py:109-131,420-427), sigma_d /
    sigma_b / mu (coeffs.py:399-410), chi2 / chi3 averaged onto the E
    positions (coeffs.py:383-384).  eps_x / eps_y are the two-point averages
    of the cell eps (coeffs.py:347-355,385).  eps_tensor = (exx, exy, eyy)
    at cell centres gives the inverse components Yee-averaged onto the E
    positions (coeffs.py:390-400); poles_tensor (fn, gamma, (sxx, sxy, syy)
    [, fc]) append one pole to each axis stack, with the off-diagonal drive
    c3xy / c3yx (coeffs.py:425-455).  Mirror cuts (y_mirror / x_mirror +-1)
    drop the low-side CPML of their axis (coeffs.py:457-469)."""
    dt, dx = grid.dt, grid.dx
    eps = _as_field(eps_inf, grid, 1.0)
    x3 = _as_field(chi3, grid, 0.0)
    x2 = _as_field(chi2, grid, 0.0)
    has_kerr = bool((x3 != 0).any())
    has_chi2 = bool((x2 != 0).any())
    cond = {}
    if eps_tensor is not None:
        # symmetric in-plane tensor at cell centres -> its inverse components
        # averaged onto the Ex / Ey positions
        exx = _as_field(eps_tensor[0], grid, 1.0)
        exy = _as_field(eps_tensor[1], grid, 0.0)
        eyy = _as_field(eps_tensor[2], grid, 1.0)
        det = exx * eyy - exy * exy
        cond["inv_xx_x"] = _avg_to_ex(eyy / det)
        cond["inv_xy_x"] = _avg_to_ex(-exy / det)
        cond["inv_yy_y"] = _avg_to_ey(exx / det)
        cond["inv_xy_y"] = _avg_to_ey(-exy / det)
    if sigma_d is not None or sigma_b is not None or mu is not None:
        sd = _as_field(sigma_d, grid, 0.0)
        sb = _as_field(sigma_b, grid, 0.0)
        mu_ = _as_field(mu, grid, 1.0)
        sdx, sdy = _avg_to_ex(sd), _avg_to_ey(sd)
        cond["d1x"] = (1 - sdx * dt / 2) / (1 + sdx * dt / 2)
        cond["d2x"] = dt / (1 + sdx * dt / 2)
        cond["d1y"] = (1 - sdy * dt / 2) / (1 + sdy * dt / 2)
        cond["d2y"] = dt / (1 + sdy * dt / 2)
        # Hz (nx-1, ny-1): four-point average of the cell centres
        sbz = 0.25 * (sb[:-1, :-1] + sb[1:, :-1] + sb[:-1, 1:] + sb[1:, 1:])
        muz = 0.25 * (mu_[:-1, :-1] + mu_[1:, :-1] + mu_[:-1, 1:] + mu_[1:, 1:])
        cond["hz1"] = (1 - sbz * dt / 2) / (1 + sbz * dt / 2)
        cond["hz2"] = dt / (muz * (1 + sbz * dt / 2))
    c1x, c2x, c3x = _axis_poles(grid, dt, eps, poles, poles_x, pole_masks, pole_masks_x,
                                (grid.nx - 1, grid.ny), _avg_to_ex)
    c1y, c2y, c3y = _axis_poles(grid, dt, eps, poles, poles_y, pole_masks, pole_masks_y,
                                (grid.nx, grid.ny - 1), _avg_to_ey)
    if len(poles_tensor):
        # oriented poles: each adds one pole to BOTH axis stacks (after the
        # scalar ones) with c3 = v3 sxx / syy times the averaged mask, and an
        # off-diagonal c3xy / c3yx = v3 sxy times the mask (zero for the
        # scalar poles) driving Px from Ey and Py from Ex
        npx0, npy0, nt = c1x.shape[0], c1y.shape[0], len(poles_tensor)
        tc1x, tc2x, tc3x, tc3xy = (grid.zeros(nt, grid.nx - 1, grid.ny) for _ in range(4))
        tc1y, tc2y, tc3y, tc3yx = (grid.zeros(nt, grid.nx, grid.ny - 1) for _ in range(4))
        for k, pole in enumerate(poles_tensor):
            fn, gamma, (sxx, sxy, syy) = pole[0], pole[1], pole[2]
            fc = pole[3] if len(pole) > 3 else fn
            v1, v2, v3 = pole_recurrence((fn, gamma, 1.0, fc), dt)
            m = (pole_masks_tensor[k] if pole_masks_tensor is not None else torch.ones_like(eps)).to(eps)
            mx, my = _avg_to_ex(m), _avg_to_ey(m)
            tc1x[k], tc2x[k] = v1, v2
            tc1y[k], tc2y[k] = v1, v2
            tc3x[k], tc3xy[k] = v3 * sxx * mx, v3 * sxy * mx
            tc3y[k], tc3yx[k] = v3 * syy * my, v3 * sxy * my
        c1x, c2x, c3x = torch.cat([c1x, tc1x]), torch.cat([c2x, tc2x]), torch.cat([c3x, tc3x])
        c1y, c2y, c3y = torch.cat([c1y, tc1y]), torch.cat([c2y, tc2y]), torch.cat([c3y, tc3y])
        cond["c3xy"] = torch.cat([grid.zeros(npx0, grid.nx - 1, grid.ny), tc3xy])
        cond["c3yx"] = torch.cat([grid.zeros(npy0, grid.nx, grid.ny - 1), tc3yx])
    if y_mirror not in (0, +1, -1) or x_mirror not in (0, +1, -1):
        raise ValueError(f"mirror parity must be 0/+1/-1, got y={y_mirror} x={x_mirror}")
    eps_x, eps_y = _avg_to_ex(eps), _avg_to_ey(eps)
    px, py = _axis_sides(npml)
    if y_mirror:    # a mirror wall replaces the low-side CPML (coeffs.py:461-469)
        py = (0, py[1])
    if x_mirror:
        px = (0, px[1])
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
    return TECoeffs(
        dt=dt, dx=dx, eps_x=eps_x, eps_y=eps_y, chi3_x=_avg_to_ex(x3), chi3_y=_avg_to_ey(x3),
        has_kerr=has_kerr, chi2_x=_avg_to_ex(x2), chi2_y=_avg_to_ey(x2), has_chi2=has_chi2,
        c1x=c1x, c2x=c2x, c3x=c3x, c1y=c1y, c2y=c2y, c3y=c3y, **cond,
        b_hx=dev(b_hx)[:, None], a_hx=dev(a_hx)[:, None], k_hx=dev(k_hx)[:, None],
        b_hy=dev(b_hy)[None, :], a_hy=dev(a_hy)[None, :], k_hy=dev(k_hy)[None, :],
        b_ey=dev(b_ey)[None, :], a_ey=dev(a_ey)[None, :], k_ey=dev(k_ey)[None, :],
        b_ex=dev(b_ex)[:, None], a_ex=dev(a_ex)[:, None], k_ex=dev(k_ex)[:, None],
        y_mirror_lo=y_mirror, x_mirror_lo=x_mirror,
        shells=shells)
