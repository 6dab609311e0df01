"""2D TM/TE restatement (CPU, torch eager) -- test infrastructure only.

Operation order follows the reference so FP32 results round identically:
  profiles      meep_torch/core/pml.py:17-59
  TM coeffs     meep_torch/core/coeffs.py:87-106,134-291 (eps, npml, y_pmc,
                electric poles, chi2/chi3, sigma_d/sigma_b/mu)
  TE coeffs     meep_torch/core/coeffs.py:357-503 (eps, npml subset)
  TM step       meep_torch/core/stepper_tm.py:78-254 (no magnetic poles/mirrors)
  TE step       meep_torch/core/stepper_te.py:73-190 (no poles/sigma/tensor eps)
  sources       meep_torch/core/sources.py:26-91
  monitors      meep_torch/core/monitors.py:11-155
  run loop      meep_torch/core/run.py:10-126
"""
import math
from dataclasses import dataclass

import torch


# ------------------------------------------------------------------ profiles
def _sides(npml):
    if isinstance(npml, (tuple, list)):
        return int(npml[0]), int(npml[1])
    return int(npml), int(npml)


def cpml_profiles(n, npml, offset, dt, dx, dtype=torch.float64):
    """CFS-PML (b, a, kappa) at x = i + offset (meep_torch/core/pml.py:17-59).

    sigma = -(m+1) ln(R0) / (2 n dx) u^m with m = 3, R0 = 1e-6; alpha = 0.05 (1-u)
    inside the layer; kappa = 1; b = exp(-(sigma + alpha) dt);
    a = sigma (b - 1) / (sigma + alpha) where sigma > 0.  The high wall is at
    n - 1 + 2 offset.
    """
    m, R0, alpha_max = 3, 1e-6, 0.05
    lo, hi = _sides(npml)
    pos = torch.arange(n, dtype=dtype) + offset
    u = torch.zeros(n, dtype=dtype)
    sig = torch.zeros_like(u)
    if lo > 0:
        ul = torch.clamp((lo - pos) / lo, min=0.0)
        u = u + ul
        sig = sig + (-(m + 1) * math.log(R0) / (2.0 * lo * dx)) * ul ** m
    if hi > 0:
        wall = n - 1 + 2.0 * offset
        uh = torch.clamp((pos - (wall - hi)) / hi, min=0.0)
        u = u + uh
        sig = sig + (-(m + 1) * math.log(R0) / (2.0 * hi * dx)) * uh ** m
    u = torch.clamp(u, max=1.0)
    alpha = torch.where(u > 0, alpha_max * (1.0 - u), torch.zeros_like(u))
    kap = torch.ones_like(u)
    b = torch.exp(-(sig / kap + alpha) * dt)
    den = kap * (sig + kap * alpha)
    a = torch.where(sig > 0, sig * (b - 1.0) / den.clamp(min=1e-300), torch.zeros_like(b))
    return b, a, kap


# -------------------------------------------------------------------- coeffs
@dataclass
class Coef2D:
    dt: float
    dx: float
    eps: torch.Tensor          # TM: eps (nx, ny); TE: cell eps (nx, ny)
    eps_x: torch.Tensor = None  # TE: (nx-1, ny)
    eps_y: torch.Tensor = None  # TE: (nx, ny-1)
    y_pmc: bool = False
    # symmetry cuts at the low edges (coeffs.py:60-74,340-344): mirror parity
    # +-1, Rotate2 phase at the low-y cut, Rotate4 phase at both
    y_mirror: int = 0
    x_mirror: int = 0
    y_rot2: int = 0
    rot4: int = 0
    # profiles as broadcastable tensors (same slicing as the reference)
    b_hy: torch.Tensor = None
    a_hy: torch.Tensor = None
    b_hx: torch.Tensor = None
    a_hx: torch.Tensor = None
    b_ex: torch.Tensor = None
    a_ex: torch.Tensor = None
    b_ey: torch.Tensor = None
    a_ey: torch.Tensor = None
    # TM media (None / empty when absent): ADE poles (np, nx, ny), chi2 / chi3
    # (nx, ny), conductivity / permeability factors
    c1: torch.Tensor = None
    c2: torch.Tensor = None
    c3: torch.Tensor = None
    chi2: torch.Tensor = None
    chi3: torch.Tensor = None
    kerr: bool = False
    d1: torch.Tensor = None
    d2: torch.Tensor = None
    h1x: torch.Tensor = None
    h2x: torch.Tensor = None
    h1y: torch.Tensor = None
    h2y: torch.Tensor = None
    # TM magnetic Lorentz poles: (nm, nx, ny-1) on Hx, (nm, nx-1, ny) on Hy
    mc1x: torch.Tensor = None
    mc2x: torch.Tensor = None
    mc3x: torch.Tensor = None
    mc1y: torch.Tensor = None
    mc2y: torch.Tensor = None
    mc3y: torch.Tensor = None
    # TE media: per-axis poles (np, nx-1, ny) / (np, nx, ny-1), sigma factors
    c1x: torch.Tensor = None
    c2x: torch.Tensor = None
    c3x: torch.Tensor = None
    c1y: torch.Tensor = None
    c2y: torch.Tensor = None
    c3y: torch.Tensor = None
    d1x: torch.Tensor = None
    d2x: torch.Tensor = None
    d1y: torch.Tensor = None
    d2y: torch.Tensor = None
    hz1: torch.Tensor = None
    hz2: torch.Tensor = None
    chi2x: torch.Tensor = None
    chi3x: torch.Tensor = None
    chi2y: torch.Tensor = None
    chi3y: torch.Tensor = None
    # TE tensor eps (inverse components on the E grids) and oriented poles
    inv_xx_x: torch.Tensor = None
    inv_xy_x: torch.Tensor = None
    inv_yy_y: torch.Tensor = None
    inv_xy_y: torch.Tensor = None
    c3xy: torch.Tensor = None
    c3yx: torch.Tensor = None


def pole_recurrence(pole, dt):
    """meep_torch/core/coeffs.py:87-106: ADE scalars of one pole
    (fn, gamma, sigma[, fc]): w0 = 2 pi fn, wc = 2 pi fc (fc = fn: Lorentz),
    G = 2 pi gamma, den = 1 + G dt / 2, c1 = (2 - w0^2 dt^2) / den,
    c2 = -(1 - G dt / 2) / den, c3 = sigma wc^2 dt^2 / den."""
    fn, gamma, sigma = pole[0], pole[1], pole[2]
    fc = pole[3] if len(pole) > 3 else fn
    w0, wc, G = 2 * math.pi * fn, 2 * math.pi * fc, 2 * math.pi * gamma
    den = 1.0 + G * dt / 2.0
    return (2.0 - w0 * w0 * dt * dt) / den, -(1.0 - G * dt / 2.0) / den, sigma * wc * wc * dt * dt / den


def _tm_media(c, nx, ny, dtype, poles, pole_masks, chi2, chi3, sigma_d, sigma_b, mu):
    """meep_torch/core/coeffs.py:154-200 (electric poles, chi2 / chi3,
    sigma_d / sigma_b / mu)."""
    dt = c.dt
    full = lambda v, d: (torch.full((nx, ny), float(d if v is None else v), dtype=dtype)
                         if not torch.is_tensor(v) else v.to(dtype))
    n = len(poles)
    c.c1 = torch.zeros(n, nx, ny, dtype=dtype)
    c.c2 = torch.zeros(n, nx, ny, dtype=dtype)
    c.c3 = torch.zeros(n, nx, ny, dtype=dtype)
    for k, pole in enumerate(poles):
        v1, v2, v3 = pole_recurrence(pole, dt)
        mask = pole_masks[k] if pole_masks is not None else torch.ones(nx, ny, dtype=dtype)
        c.c1[k] = v1 * torch.ones(nx, ny, dtype=dtype)
        c.c2[k] = v2 * torch.ones(nx, ny, dtype=dtype)
        c.c3[k] = v3 * mask.to(dtype)
    c.chi2, c.chi3 = full(chi2, 0.0), full(chi3, 0.0)
    c.kerr = bool((c.chi2 != 0).any()) or bool((c.chi3 != 0).any())
    if sigma_d is not None or sigma_b is not None or mu is not None:
        sd, sb, mu_ = full(sigma_d, 0.0), full(sigma_b, 0.0), full(mu, 1.0)
        c.d1 = (1 - sd * dt / 2) / (1 + sd * dt / 2)
        c.d2 = dt / (1 + sd * dt / 2)
        sbx, mux = 0.5 * (sb[:, :-1] + sb[:, 1:]), 0.5 * (mu_[:, :-1] + mu_[:, 1:])
        sby, muy = 0.5 * (sb[:-1, :] + sb[1:, :]), 0.5 * (mu_[:-1, :] + mu_[1:, :])
        c.h1x = (1 - sbx * dt / 2) / (1 + sbx * dt / 2)
        c.h2x = dt / (mux * (1 + sbx * dt / 2))
        c.h1y = (1 - sby * dt / 2) / (1 + sby * dt / 2)
        c.h2y = dt / (muy * (1 + sby * dt / 2))
    return c


def _tm_mpoles(c, nx, ny, dtype, m_poles, m_pole_masks):
    """meep_torch/core/coeffs.py:177-196: per magnetic pole (fn, gamma,
    sigma), w0 = 2 pi fn, G = 2 pi gamma, den = 1 + G dt / 2:
    mc1 = (2 - w0^2 dt^2) / den, mc2 = -(1 - G dt / 2) / den and
    mc3 = (sigma w0^2 dt^2 / den) * (mask averaged onto the H component)."""
    dt, nm = c.dt, len(m_poles)
    for suf, shape, avg in (("x", (nx, ny - 1), lambda v: 0.5 * (v[:, :-1] + v[:, 1:])),
                            ("y", (nx - 1, ny), lambda v: 0.5 * (v[:-1, :] + v[1:, :]))):
        m1, m2, m3 = (torch.zeros(nm, *shape, dtype=dtype) for _ in range(3))
        for k, (fn, gamma, sigma) in enumerate(p[:3] for p in m_poles):
            w0, G = 2 * math.pi * fn, 2 * math.pi * gamma
            den = 1.0 + G * dt / 2.0
            mask = m_pole_masks[k] if m_pole_masks is not None else torch.ones(nx, ny, dtype=dtype)
            m1[k] = (2.0 - w0 * w0 * dt * dt) / den
            m2[k] = -(1.0 - G * dt / 2.0) / den
            m3[k] = (sigma * w0 * w0 * dt * dt / den) * avg(mask.to(dtype))
        setattr(c, "mc1" + suf, m1)
        setattr(c, "mc2" + suf, m2)
        setattr(c, "mc3" + suf, m3)


def tm_coeffs(nx, ny, dx, eps, npml, courant=0.5, dtype=torch.float32, y_pmc=False, poles=(),
              pole_masks=None, chi2=None, chi3=None, sigma_d=None, sigma_b=None, mu=None, m_poles=(),
              m_pole_masks=None, y_mirror=0, x_mirror=0, y_rot2=0, rot4=0):
    """meep_torch/core/coeffs.py:134-291: eps + CPML (+ y_pmc), electric ADE
    poles, chi2 / chi3, sigma_d / sigma_b / mu and magnetic Lorentz poles
    (not combined with sigma_b / mu, as in the reference).  Symmetry cuts
    (coeffs.py:211-281): a cut at the low edge of an axis drops that side's
    CPML, and the D-update profile of that axis keeps nodes 0..n-2 (the cut
    node is live) instead of the interior 1..n-2."""
    dt = courant * dx
    eps = eps.to(dtype)
    if isinstance(npml, (tuple, list)):
        px, py = _sides(npml[0]), _sides(npml[1])
    else:
        px = py = (int(npml), int(npml))
    ycut, xcut = bool(y_mirror or y_rot2 or rot4), bool(x_mirror or rot4)
    if ycut:
        py = (0, py[1])
    if xcut:
        px = (0, px[1])

    def prof(n, off, pair):
        return cpml_profiles(n, pair, off, dt, dx, dtype)

    b_hy, a_hy, _ = prof(ny - 1, 0.5, (0, 0) if y_pmc else py)
    b_hx, a_hx, _ = prof(nx - 1, 0.5, px)
    b_exf, a_exf, _ = prof(nx, 0.0, px)
    b_eyf, a_eyf, _ = prof(ny, 0.0, (0, 0) if y_pmc else py)
    if y_pmc:
        b_ey, a_ey = b_eyf, a_eyf
    elif ycut:
        b_ey, a_ey = b_eyf[:-1], a_eyf[:-1]
    else:
        b_ey, a_ey = b_eyf[1:-1], a_eyf[1:-1]
    b_ex, a_ex = (b_exf[:-1], a_exf[:-1]) if xcut else (b_exf[1:-1], a_exf[1:-1])
    c = Coef2D(dt=dt, dx=dx, eps=eps, y_pmc=y_pmc, y_mirror=y_mirror, x_mirror=x_mirror, y_rot2=y_rot2,
               rot4=rot4, b_hy=b_hy[None, :], a_hy=a_hy[None, :],
               b_hx=b_hx[:, None], a_hx=a_hx[:, None],
               b_ex=b_ex[:, None], a_ex=a_ex[:, None],
               b_ey=b_ey[None, :], a_ey=a_ey[None, :])
    c = _tm_media(c, nx, ny, dtype, poles, pole_masks, chi2, chi3, sigma_d, sigma_b, mu)
    if len(m_poles):
        if c.h1x is not None:
            raise NotImplementedError("m_poles with sigma_b/mu")
        _tm_mpoles(c, nx, ny, dtype, m_poles, m_pole_masks)
    return c


def _te_axis_poles(nx, ny, dt, dtype, shared, own, shared_masks, own_masks, shape, avg):
    """meep_torch/core/coeffs.py:109-131: shared poles first, then the
    axis' own; each cell mask (ones if absent) Yee-averaged by ``avg``."""
    allp = list(shared) + list(own)
    c1, c2, c3 = (torch.zeros(len(allp), *shape, dtype=dtype) for _ in range(3))
    for k, pole in enumerate(allp):
        v1, v2, v3 = pole_recurrence(pole, dt)
        if k < len(shared):
            m = shared_masks[k] if shared_masks is not None else None
        else:
            m = own_masks[k - len(shared)] if own_masks is not None else None
        mask = (m if m is not None else torch.ones(nx, ny, dtype=dtype)).to(dtype)
        c1[k] = v1
        c2[k] = v2
        c3[k] = v3 * avg(mask)
    return c1, c2, c3


def te_coeffs(nx, ny, dx, eps, npml, courant=0.5, dtype=torch.float32, poles=(), pole_masks=None, poles_x=(),
              poles_y=(), pole_masks_x=None, pole_masks_y=None, sigma_d=None, sigma_b=None, mu=None, chi2=None,
              chi3=None, eps_tensor=None, poles_tensor=(), pole_masks_tensor=None, y_mirror=0, x_mirror=0):
    """meep_torch/core/coeffs.py:357-503: eps + CPML, shared and per-axis
    electric poles, sigma_d / sigma_b / mu, chi2 / chi3, the in-plane tensor
    eps (exx, exy, eyy) -> inverse components averaged onto Ex / Ey
    (coeffs.py:390-400) and oriented poles (fn, gamma, (sxx, sxy, syy)[, fc])
    appended to both axis stacks with the off-diagonal c3xy / c3yx
    (coeffs.py:425-455); eps_x/eps_y are the two-point averages of
    coeffs.py:347-355.  A mirror cut at the low edge of an axis drops that
    side's CPML (coeffs.py:457-469)."""
    dt = courant * dx
    eps = eps.to(dtype)
    eps_x = 0.5 * (eps[:-1, :] + eps[1:, :])
    eps_y = 0.5 * (eps[:, :-1] + eps[:, 1:])
    if isinstance(npml, (tuple, list)):
        px, py = _sides(npml[0]), _sides(npml[1])
    else:
        px = py = (int(npml), int(npml))
    if y_mirror:
        py = (0, py[1])
    if x_mirror:
        px = (0, px[1])

    def prof(n, off, pair):
        return cpml_profiles(n, pair, off, dt, dx, dtype)

    b_hx, a_hx, _ = prof(nx - 1, 0.5, px)
    b_hy, a_hy, _ = prof(ny - 1, 0.5, py)
    b_eyf, a_eyf, _ = prof(ny, 0.0, py)
    b_exf, a_exf, _ = prof(nx, 0.0, px)
    c = Coef2D(dt=dt, dx=dx, eps=eps, eps_x=eps_x, eps_y=eps_y, y_mirror=y_mirror, x_mirror=x_mirror,
               b_hx=b_hx[:, None], a_hx=a_hx[:, None],
               b_hy=b_hy[None, :], a_hy=a_hy[None, :],
               b_ey=b_eyf[1:-1][None, :], a_ey=a_eyf[1:-1][None, :],
               b_ex=b_exf[1:-1][:, None], a_ex=a_exf[1:-1][:, None])
    ax = lambda v: 0.5 * (v[:-1, :] + v[1:, :])
    ay = lambda v: 0.5 * (v[:, :-1] + v[:, 1:])
    c.c1x, c.c2x, c.c3x = _te_axis_poles(nx, ny, dt, dtype, poles, poles_x, pole_masks, pole_masks_x,
                                         (nx - 1, ny), ax)
    c.c1y, c.c2y, c.c3y = _te_axis_poles(nx, ny, dt, dtype, poles, poles_y, pole_masks, pole_masks_y,
                                         (nx, ny - 1), ay)
    if eps_tensor is not None:
        fl = lambda v, d: (torch.full((nx, ny), float(d if v is None else v), dtype=dtype)
                           if not torch.is_tensor(v) else v.to(dtype))
        exx, exy, eyy = fl(eps_tensor[0], 1.0), fl(eps_tensor[1], 0.0), fl(eps_tensor[2], 1.0)
        det = exx * eyy - exy * exy
        c.inv_xx_x, c.inv_xy_x = ax(eyy / det), ax(-exy / det)
        c.inv_yy_y, c.inv_xy_y = ay(exx / det), ay(-exy / det)
    if len(poles_tensor):
        nt, npx0, npy0 = len(poles_tensor), c.c1x.shape[0], c.c1y.shape[0]
        t1x, t2x, t3x, t3xy = (torch.zeros(nt, nx - 1, ny, dtype=dtype) for _ in range(4))
        t1y, t2y, t3y, t3yx = (torch.zeros(nt, nx, ny - 1, dtype=dtype) for _ in range(4))
        for k, pole in enumerate(poles_tensor):
            fn, gamma, (sxx, sxy, syy) = pole[0], pole[1], pole[2]
            v1, v2, v3 = pole_recurrence((fn, gamma, 1.0, pole[3] if len(pole) > 3 else fn), dt)
            m = (pole_masks_tensor[k] if pole_masks_tensor is not None else torch.ones(nx, ny)).to(dtype)
            mx, my = ax(m), ay(m)
            t1x[k], t2x[k], t1y[k], t2y[k] = v1, v2, v1, v2
            t3x[k], t3xy[k] = v3 * sxx * mx, v3 * sxy * mx
            t3y[k], t3yx[k] = v3 * syy * my, v3 * sxy * my
        c.c1x, c.c2x, c.c3x = torch.cat([c.c1x, t1x]), torch.cat([c.c2x, t2x]), torch.cat([c.c3x, t3x])
        c.c1y, c.c2y, c.c3y = torch.cat([c.c1y, t1y]), torch.cat([c.c2y, t2y]), torch.cat([c.c3y, t3y])
        c.c3xy = torch.cat([torch.zeros(npx0, nx - 1, ny, dtype=dtype), t3xy])
        c.c3yx = torch.cat([torch.zeros(npy0, nx, ny - 1, dtype=dtype), t3yx])
    full = lambda v: (torch.full((nx, ny), float(v or 0.0), dtype=dtype) if not torch.is_tensor(v) else v.to(dtype))
    x2, x3 = full(chi2), full(chi3)
    c.kerr = bool((x2 != 0).any()) or bool((x3 != 0).any())   # coeffs.py:380-384
    c.chi2x, c.chi2y, c.chi3x, c.chi3y = ax(x2), ay(x2), ax(x3), ay(x3)
    if sigma_d is not None or sigma_b is not None or mu is not None:
        # coeffs.py:399-410
        full = lambda v, d: (torch.full((nx, ny), float(d if v is None else v), dtype=dtype)
                             if not torch.is_tensor(v) else v.to(dtype))
        sd, sb, mu_ = full(sigma_d, 0.0), full(sigma_b, 0.0), full(mu, 1.0)
        sdx, sdy = ax(sd), ay(sd)
        c.d1x, c.d2x = (1 - sdx * dt / 2) / (1 + sdx * dt / 2), dt / (1 + sdx * dt / 2)
        c.d1y, c.d2y = (1 - sdy * dt / 2) / (1 + sdy * dt / 2), dt / (1 + sdy * dt / 2)
        sbz = 0.25 * (sb[:-1, :-1] + sb[1:, :-1] + sb[:-1, 1:] + sb[1:, 1:])
        muz = 0.25 * (mu_[:-1, :-1] + mu_[1:, :-1] + mu_[:-1, 1:] + mu_[1:, 1:])
        c.hz1 = (1 - sbz * dt / 2) / (1 + sbz * dt / 2)
        c.hz2 = dt / (muz * (1 + sbz * dt / 2))
    return c


# --------------------------------------------------------------------- state
def tm_state(nx, ny, dtype=torch.float32, y_pmc=False, n_poles=0, n_mpoles=0, y_mirror=0, x_mirror=0, y_rot2=0,
             rot4=0):
    """Field dict with the TMState shapes of stepper_tm.py:12-71 (pole
    stacks only when n_poles > 0; B and the magnetic pole stacks only when
    n_mpoles > 0; the on-plane H rows Hx0 (nx,) under an odd y mirror,
    Rotate2 or Rotate4 and Hy0 (ny,) under an odd x mirror or Rotate4; a cut
    makes the D-update psi one node longer on its axis)."""
    z = lambda *s: torch.zeros(*s, dtype=dtype)
    ny_e = ny if y_pmc else (ny - 1 if (y_mirror or y_rot2 or rot4) else ny - 2)
    nx_e = nx - 1 if (x_mirror or rot4) else nx - 2
    out = {"Ez": z(nx, ny), "Hx": z(nx, ny - 1), "Hy": z(nx - 1, ny), "Dz": z(nx, ny),
           "psi_Ezy": z(nx, ny - 1), "psi_Ezx": z(nx - 1, ny),
           "psi_Hyx": z(nx_e, ny_e), "psi_Hxy": z(nx_e, ny_e)}
    if y_mirror == -1 or y_rot2 or rot4:
        out["Hx0"] = z(nx)
    if x_mirror == -1 or rot4:
        out["Hy0"] = z(ny)
    if n_poles:
        out.update(P_cur=z(n_poles, nx, ny), P_prev=z(n_poles, nx, ny))
    if n_mpoles:
        out.update(Bx=z(nx, ny - 1), By=z(nx - 1, ny), Mx_cur=z(n_mpoles, nx, ny - 1),
                   Mx_prev=z(n_mpoles, nx, ny - 1), My_cur=z(n_mpoles, nx - 1, ny),
                   My_prev=z(n_mpoles, nx - 1, ny))
    return out


def te_state(nx, ny, dtype=torch.float32, npx=0, npy=0):
    """Field dict with the TEState shapes of stepper_te.py:11-49 (pole stacks
    only when present)."""
    z = lambda *s: torch.zeros(*s, dtype=dtype)
    out = {"Hz": z(nx - 1, ny - 1), "Ex": z(nx - 1, ny), "Ey": z(nx, ny - 1),
           "Dx": z(nx - 1, ny), "Dy": z(nx, ny - 1),
           "psi_Eyx": z(nx - 1, ny - 1), "psi_Exy": z(nx - 1, ny - 1),
           "psi_Hzy": z(nx - 1, ny - 2), "psi_Hzx": z(nx - 2, ny - 1)}
    if npx:
        out.update(Px_cur=z(npx, nx - 1, ny), Px_prev=z(npx, nx - 1, ny))
    if npy:
        out.update(Py_cur=z(npy, nx, ny - 1), Py_prev=z(npy, nx, ny - 1))
    return out


# ---------------------------------------------------------------------- steps
def tm_step(s, c, coef=0.0, mask=None, hcoef=0.0, hx_mask=None, hy_mask=None):
    """One TM step: B half (stepper_tm.py:85-92,126-128), the TFSF magnetic
    currents of the B phase (stepper_tm.py:180-193: H -= (dt*h_amp) * mask),
    D half (stepper_tm.py:134-165, PEC or y_pmc), soft source
    (stepper_tm.py:197-200: Dz -= (dt*amp) * mask, the scalar formed in
    double first), the ADE poles and Ez = (Dz - sum P) / eps with the chi2 /
    chi3 fixed point (stepper_tm.py:203-231).  ``coef`` is dt*amp (the raw
    amp under sigma_d, scaled by d2), ``hcoef`` dt*h_amp."""
    dx, dt = c.dx, c.dt
    Ez = s["Ez"]
    g = (Ez[:, 1:] - Ez[:, :-1]) / dx
    pzy = c.b_hy * s["psi_Ezy"] + c.a_hy * g
    g2 = (Ez[1:, :] - Ez[:-1, :]) / dx
    pzx = c.b_hx * s["psi_Ezx"] + c.a_hx * g2
    mag = {}
    # on-plane H rows of the odd cuts (stepper_tm.py:94-115): a ghost Ez
    # beyond the cut gives dEz/dn = (Ez_cut - ghost) / dx; the update is
    # H0 -/+ (dt * (Ez_cut - ghost)) / dx with Ez at time n
    if "Hx0" in s:
        ez0 = Ez[:, 0]
        ghost = c.rot4 * Ez[0, :] if c.rot4 else (c.y_rot2 * ez0.flip(0) if c.y_rot2 else -ez0)
        mag["Hx0"] = s["Hx0"] - dt * (ez0 - ghost) / dx
    if "Hy0" in s:
        ez0x = Ez[0, :]
        ghost_x = c.rot4 * Ez[:, 0] if c.rot4 else -ez0x
        mag["Hy0"] = s["Hy0"] + dt * (ez0x - ghost_x) / dx
    if c.mc1x is not None:      # stepper_tm.py:103-106: B integrates the curl
        Bx = s["Bx"] - dt * (g + pzy)
        By = s["By"] + dt * (g2 + pzx)
        if hy_mask is not None:     # stepper_tm.py:185-193: the TFSF currents go into B
            By = By - hcoef * hy_mask
        if hx_mask is not None:
            Bx = Bx - hcoef * hx_mask
        # stepper_tm.py:214-224: M' = mc1 M + mc2 M_prev + mc3 H(n-1/2), H = B - sum M'
        Mx = c.mc1x * s["Mx_cur"] + c.mc2x * s["Mx_prev"] + c.mc3x * s["Hx"]
        My = c.mc1y * s["My_cur"] + c.mc2y * s["My_prev"] + c.mc3y * s["Hy"]
        Hx, Hy = Bx - Mx.sum(0), By - My.sum(0)
        mag.update(Bx=Bx, By=By, Mx_cur=Mx, Mx_prev=s["Mx_cur"], My_cur=My, My_prev=s["My_cur"])
    else:
        if c.h1x is not None:       # stepper_tm.py:116-121: dB/dt = -curl E - sigma_b B, mu folded in h2
            Hx = c.h1x * s["Hx"] - c.h2x * (g + pzy)
            Hy = c.h1y * s["Hy"] + c.h2y * (g2 + pzx)
        else:
            Hx = s["Hx"] - dt * (g + pzy)
            Hy = s["Hy"] + dt * (g2 + pzx)
        if hy_mask is not None:
            Hy = Hy - hcoef * hy_mask
        if hx_mask is not None:
            Hx = Hx - hcoef * hx_mask

    # stepper_tm.py:134-155: the H beyond each wall (PMC: zero both sides;
    # even mirror: a zero ghost at the cut; odd / Rotate cuts: the on-plane
    # row), then differences over the live Ez region
    if c.y_pmc:
        Hxp = torch.nn.functional.pad(Hx, (1, 1))
        ys, pad_y = slice(None), (0, 0)
    elif c.y_mirror == 1:
        Hxp, ys, pad_y = torch.nn.functional.pad(Hx, (1, 0)), slice(None, -1), (0, 1)
    elif c.y_mirror == -1 or c.y_rot2 or c.rot4:
        Hxp, ys, pad_y = torch.cat([mag["Hx0"][:, None], Hx], dim=1), slice(None, -1), (0, 1)
    else:
        Hxp, ys, pad_y = Hx, slice(1, -1), (1, 1)
    if c.x_mirror == 1:
        Hyp, xs, pad_x = torch.nn.functional.pad(Hy, (0, 0, 1, 0)), slice(None, -1), (0, 1)
    elif c.x_mirror == -1 or c.rot4:
        Hyp, xs, pad_x = torch.cat([mag["Hy0"][None, :], Hy], dim=0), slice(None, -1), (0, 1)
    else:
        Hyp, xs, pad_x = Hy, slice(1, -1), (1, 1)
    dHx = (Hxp[xs, 1:] - Hxp[xs, :-1]) / dx
    dHy = (Hyp[1:, ys] - Hyp[:-1, ys]) / dx
    pyx = c.b_ex * s["psi_Hyx"] + c.a_ex * dHy
    pxy = c.b_ey * s["psi_Hxy"] + c.a_ey * dHx
    curl = dHy + pyx - dHx - pxy
    if c.d1 is not None:        # stepper_tm.py:161-163: dD/dt = curl H - sigma_d D
        Dz = c.d1 * s["Dz"] + c.d2 * torch.nn.functional.pad(curl, pad_y + pad_x)
    else:
        Dz = s["Dz"] + dt * torch.nn.functional.pad(curl, pad_y + pad_x)
    if mask is not None:
        if c.d1 is not None:    # stepper_tm.py:198-200: scale = d2, coef is then the raw amplitude
            Dz = Dz - c.d2 * coef * mask
        else:
            Dz = Dz - coef * mask
    out = dict(s)
    out.update(mag)
    rhs = Dz
    if c.c1 is not None and c.c1.shape[0]:
        # stepper_tm.py:203-208: P(n+1) = c1 P(n) + c2 P(n-1) + c3 Ez(n)
        P_new = c.c1 * s["P_cur"] + c.c2 * s["P_prev"] + c.c3 * Ez
        out.update(P_cur=P_new, P_prev=s["P_cur"])
        rhs = Dz - P_new.sum(0)
    # stepper_tm.py:226-231: E = rhs / eps, then 3 fixed-point passes of
    # E = rhs / (eps + chi2 E + chi3 E^2) when chi2 or chi3 is present
    Ez_new = rhs / c.eps
    if c.kerr:
        for _ in range(3):
            Ez_new = rhs / (c.eps + c.chi2 * Ez_new + c.chi3 * Ez_new * Ez_new)
    out.update(Hx=Hx, Hy=Hy, Dz=Dz, Ez=Ez_new, psi_Ezy=pzy, psi_Ezx=pzx,
               psi_Hyx=pyx, psi_Hxy=pxy)
    return out


def te_step(s, c, coef=0.0, mask=None, hcoef=0.0, hz_mask=None):
    """One TE step entered at the D half (stepper_te.py:178-190): D from curl
    Hz (stepper_te.py:80-103), source on Dy (stepper_te.py:131-134), E = D/eps
    (stepper_te.py:160-168), then Hz from curl E (stepper_te.py:107-117) and
    the TFSF magnetic current on Hz (stepper_te.py:126-128)."""
    dx, dt = c.dx, c.dt
    Hz = s["Hz"]
    gy = (Hz[:, 1:] - Hz[:, :-1]) / dx
    phy = c.b_ey * s["psi_Hzy"] + c.a_ey * gy
    cx = torch.nn.functional.pad(gy + phy, (1, 1))
    gx = (Hz[1:, :] - Hz[:-1, :]) / dx
    phx = c.b_ex * s["psi_Hzx"] + c.a_ex * gx
    cy = torch.nn.functional.pad(gx + phx, (0, 0, 1, 1))
    # even mirror (stepper_te.py:88-97): the on-plane tangential E row is
    # live, with the odd Hz ghost giving dHz/dn = 2 Hz[0] / dx there
    if c.y_mirror == 1:
        cx = torch.cat([2.0 * Hz[:, :1] / dx, cx[:, 1:]], dim=1)
    if c.x_mirror == 1:
        cy = torch.cat([2.0 * Hz[:1, :] / dx, cy[1:, :]], dim=0)
    if c.d1x is not None:       # stepper_te.py:96-99 (sigma_d)
        Dx = c.d1x * s["Dx"] + c.d2x * cx
        Dy = c.d1y * s["Dy"] - c.d2y * cy
    else:
        Dx = s["Dx"] + dt * cx
        Dy = s["Dy"] - dt * cy
    if mask is not None:
        # stepper_te.py:131-134: scale d2y under sigma_d (coef is then the raw amplitude)
        Dy = Dy - (c.d2y * coef * mask if c.d1x is not None else coef * mask)
    out = dict(s)
    rx, ry = Dx, Dy
    # four-point averages of the other component, zeros beyond the walls (stepper_te.py:52-61)
    a4x = lambda v: (lambda q: 0.25 * (q[:-1, :-1] + q[:-1, 1:] + q[1:, :-1] + q[1:, 1:]))(
        torch.nn.functional.pad(v, (1, 1)))
    a4y = lambda v: (lambda q: 0.25 * (q[:-1, :-1] + q[:-1, 1:] + q[1:, :-1] + q[1:, 1:]))(
        torch.nn.functional.pad(v, (0, 0, 1, 1)))
    if c.c1x is not None and c.c1x.shape[0]:   # stepper_te.py:137-147
        Pxn = c.c1x * s["Px_cur"] + c.c2x * s["Px_prev"] + c.c3x * s["Ex"]
        if c.c3xy is not None:                  # oriented poles: Px driven by E(n-1/2) of the other axis
            Pxn = Pxn + c.c3xy * a4x(s["Ey"])
        out.update(Px_cur=Pxn, Px_prev=s["Px_cur"])
        rx = Dx - Pxn.sum(0)
    if c.c1y is not None and c.c1y.shape[0]:
        Pyn = c.c1y * s["Py_cur"] + c.c2y * s["Py_prev"] + c.c3y * s["Ey"]
        if c.c3yx is not None:
            Pyn = Pyn + c.c3yx * a4y(s["Ex"])
        out.update(Py_cur=Pyn, Py_prev=s["Py_cur"])
        ry = Dy - Pyn.sum(0)
    if c.inv_xx_x is not None:                  # stepper_te.py:158-160: E = eps^-1 (D - P)
        Ex = c.inv_xx_x * rx + c.inv_xy_x * a4x(ry)
        Ey = c.inv_yy_y * ry + c.inv_xy_y * a4y(rx)
    else:
        Ex = rx / c.eps_x
        Ey = ry / c.eps_y
    if c.kerr:
        # stepper_te.py:161-175: Jacobi passes, |E|^2 with the four-point
        # average of the other component, zero-padded beyond the walls
        # (stepper_te.py:52-61)
        for _ in range(3):
            eyp = torch.nn.functional.pad(Ey, (1, 1))
            exp_ = torch.nn.functional.pad(Ex, (0, 0, 1, 1))
            ay4 = 0.25 * (eyp[:-1, :-1] + eyp[:-1, 1:] + eyp[1:, :-1] + eyp[1:, 1:])
            ax4 = 0.25 * (exp_[:-1, :-1] + exp_[:-1, 1:] + exp_[1:, :-1] + exp_[1:, 1:])
            e2x = Ex * Ex + ay4 ** 2
            e2y = Ey * Ey + ax4 ** 2
            Ex = rx / (c.eps_x + c.chi2x * Ex + c.chi3x * e2x)
            Ey = ry / (c.eps_y + c.chi2y * Ey + c.chi3y * e2y)
    dEy = (Ey[1:, :] - Ey[:-1, :]) / dx
    dEx = (Ex[:, 1:] - Ex[:, :-1]) / dx
    pyx = c.b_hx * s["psi_Eyx"] + c.a_hx * dEy
    pxy = c.b_hy * s["psi_Exy"] + c.a_hy * dEx
    if c.hz1 is not None:       # stepper_te.py:110-115 (sigma_b, mu)
        Hz2 = c.hz1 * Hz - c.hz2 * (dEy + pyx - dEx - pxy)
    else:
        Hz2 = Hz - dt * (dEy + pyx - dEx - pxy)
    if hz_mask is not None:
        Hz2 = Hz2 - hcoef * hz_mask
    out.update(Dx=Dx, Dy=Dy, Ex=Ex, Ey=Ey, Hz=Hz2, psi_Hzy=phy, psi_Hzx=phx,
               psi_Eyx=pyx, psi_Exy=pxy)
    return out


# -------------------------------------------------------------------- sources
def gaussian_pulse(fcen, fwidth, cutoff=5.0):
    """meep_torch/core/sources.py:26-36: t -> cos(2 pi f (t-t0)) exp(-(t-t0)^2/(2w^2)),
    w = 1/fwidth, t0 = cutoff*w."""
    w = 1.0 / fwidth
    t0 = cutoff * w

    def amp(t):
        return math.cos(2 * math.pi * fcen * (t - t0)) * math.exp(-((t - t0) ** 2) / (2 * w ** 2))
    return amp


def line_mask(nx, ny, ix, j0, j1, dx, dtype=torch.float32, weights=None):
    """meep_torch/core/sources.py:49-78: unit (or per-cell) weights on
    x = ix, j in [j0, j1), scaled by dx/dx^2 for a segment and 1/dx^2 for a
    single cell."""
    m = torch.zeros(nx, ny, dtype=dtype)
    if weights is None:
        m[ix, j0:j1] = 1.0
    else:
        m[ix, j0:j1] = torch.as_tensor(weights, dtype=dtype)
    per = dx if (j1 - j0) > 1 else 1.0
    return m * (per / dx ** 2)


# ------------------------------------------------------------------ monitors
def _phasor(freqs, dtype):
    """2*pi*f as a grid-dtype tensor (monitors.py:65,86)."""
    return 2 * math.pi * torch.tensor(freqs, dtype=dtype)


def flux_line_sample(kind, s, idx, lo, hi, freqs, dt, t_idx, dtype):
    """Per-step (F, 4, L) DFT sample of a flux line (monitors.py:71-95,
    113-122, 141-151): dt * [cos(w tE) E, -sin(w tE) E, cos(w tH) H, -sin(w tH) H].

    kind: 'tm_x' | 'tm_y' | 'te_x' | 'te_y'.
    """
    if kind == "tm_x":
        E = s["Ez"][idx, lo:hi]
        H = 0.5 * (s["Hy"][idx - 1, lo:hi] + s["Hy"][idx, lo:hi])
        tE, tH = (t_idx + 1.0) * dt, (t_idx + 0.5) * dt
    elif kind == "tm_y":
        E = s["Ez"][lo:hi, idx]
        H = 0.5 * (s["Hx"][lo:hi, idx - 1] + s["Hx"][lo:hi, idx])
        tE, tH = (t_idx + 1.0) * dt, (t_idx + 0.5) * dt
    elif kind == "te_x":
        E = s["Ey"][idx, lo:hi]
        H = 0.5 * (s["Hz"][idx - 1, lo:hi] + s["Hz"][idx, lo:hi])
        tE, tH = (t_idx + 0.5) * dt, (t_idx + 1.0) * dt
    else:
        E = s["Ex"][lo:hi, idx]
        H = 0.5 * (s["Hz"][lo:hi, idx - 1] + s["Hz"][lo:hi, idx])
        tE, tH = (t_idx + 0.5) * dt, (t_idx + 1.0) * dt
    w = _phasor(freqs, dtype)[:, None]
    return dt * torch.stack([torch.cos(w * tE) * E, -torch.sin(w * tE) * E,
                             torch.cos(w * tH) * H, -torch.sin(w * tH) * H], dim=1)


def flux_from_acc(kind, acc, dx):
    """monitors.py:97-100,124-126,153-155: -/+ 1/2 sum(ReE ReH + ImE ImH) dx."""
    ReE, ImE, ReH, ImH = acc.unbind(1)
    s = (ReE * ReH + ImE * ImH).sum(-1) * dx
    return {"tm_x": -0.5, "tm_y": 0.5, "te_x": 0.5, "te_y": -0.5}[kind] * s


def dft_field_sample(field, comp, freqs, dt, t_idx, dtype):
    """monitors.py:40-46: dt * [cos(w t) F, -sin(w t) F] with t = (t+1)dt for
    E components, (t+1/2)dt for H components (TM timing on TE runs too)."""
    off = 1.0 if comp.startswith("E") else 0.5
    t = (t_idx + off) * dt
    w = _phasor(freqs, dtype)[:, None, None]
    return dt * torch.stack([torch.cos(w * t) * field, -torch.sin(w * t) * field], dim=1)


# ------------------------------------------------------------------- runner
def run2d(mode, state, c, n_steps, amps=None, mask=None, monitors=(), checkpoint_every=0, hamps=None,
          hmasks=None):
    """meep_torch/core/run.py:10-63,86-97: chunked forward run.  Per chunk the
    monitor samples are summed in step order starting from the first sample;
    chunk sums are added to the running total.  ``amps`` are the per-step
    source amplitudes (the profile at (t+1/2)dt, sources.py:90-91).

    monitors: list of dicts {"type": "flux", "kind", "idx", "lo", "hi", "freqs"}
    or {"type": "dft", "comp", "freqs"} or {"type": "probe", "ix", "iy"}.
    ``hamps`` / ``hmasks`` (optional): per-step magnetic amplitudes and the
    TFSF masks {"hx_mask", "hy_mask"} (TM) or {"hz_mask"} (TE), fed to the
    step as hcoef = dt * hamps[t] (run.py:17-23).
    """
    step = tm_step if mode == "tm" else te_step
    dtype = state[next(iter(state))].dtype
    acc = [None] * len(monitors)
    chunk = n_steps if checkpoint_every <= 0 else checkpoint_every
    t0 = 0
    s = state
    while t0 < n_steps:
        k = min(chunk, n_steps - t0)
        part = [None] * len(monitors)
        for t in range(t0, t0 + k):
            hk = {} if hamps is None else dict(hmasks, hcoef=c.dt * hamps[t])
            if amps is not None:
                # under sigma_d the source is scaled by d2 instead of dt (stepper_tm.py:198-200)
                raw = getattr(c, "d1", None) is not None or getattr(c, "d1x", None) is not None
                a = amps[t] if raw else c.dt * amps[t]
                s = step(s, c, a, mask, **hk)
            else:
                s = step(s, c, **hk)
            for i, mon in enumerate(monitors):
                if mon["type"] == "flux":
                    smp = flux_line_sample(mon["kind"], s, mon["idx"], mon["lo"], mon["hi"],
                                           mon["freqs"], c.dt, t, dtype)
                elif mon["type"] == "dft":
                    smp = dft_field_sample(s[mon["comp"]], mon["comp"], mon["freqs"], c.dt, t, dtype)
                else:
                    smp = s["Ez"][mon["ix"], mon["iy"]].clone()[None]
                if mon["type"] == "probe":
                    part[i] = smp if part[i] is None else torch.cat([part[i], smp])
                else:
                    part[i] = smp if part[i] is None else part[i] + smp
        for i, mon in enumerate(monitors):
            if acc[i] is None:
                acc[i] = part[i]
            elif mon["type"] == "probe":
                acc[i] = torch.cat([acc[i], part[i]])
            else:
                acc[i] = acc[i] + part[i]
        t0 += k
    return s, acc


# ---------------------------------------------------------------- invariants
def tm_yee_invariant(s_prev, s_next, eps, dx):
    """Discrete TM energy conserved by the leapfrog in a lossless PEC box
    (tests/analytic/test_energy.py:29-41): 1/2 sum eps Ez(n)^2 +
    1/2 sum Hx(n-1/2) Hx(n+1/2) + 1/2 sum Hy(n-1/2) Hy(n+1/2)."""
    e = 0.5 * (eps * s_prev["Ez"] ** 2).sum()
    h = 0.5 * (s_prev["Hx"] * s_next["Hx"]).sum() + 0.5 * (s_prev["Hy"] * s_next["Hy"]).sum()
    return float((e + h) * dx * dx)
