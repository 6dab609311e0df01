"""3D Yee restatement (CPU, torch eager) -- test infrastructure only.

Follows meep_torch/core/stepper_3d.py:150-266 (CPML coefficients, the
matched-sigma absorber and mixed boundary modes; material sigma_d / sigma_b
damping pairs and ADE poles, stepper_3d.py:217-266), pml.py:62-80 (absorber
ramp), stepper_3d.py:295-301,328-361,363-445 (absorber and CPML steps,
folded-D pole update),
stepper_3d.py:448-481 (sources),
stepper_3d_lean.py:174-296 (in-place lean form) and monitors_3d.py:18-210.
"""
import math

import torch

from .yee2d import _sides, cpml_profiles, pole_recurrence


def avg(v, ax):
    """Two-point average along ``ax`` (stepper_3d.py:70-75)."""
    n = v.shape[ax]
    return 0.5 * (v.narrow(ax, 0, n - 1) + v.narrow(ax, 1, n - 1))


def diff(v, ax):
    """Forward difference along ``ax`` (stepper_3d.py:78-83)."""
    n = v.shape[ax]
    return v.narrow(ax, 1, n - 1) - v.narrow(ax, 0, n - 1)


def _along(v, ax):
    sh = [1, 1, 1]
    sh[ax] = v.shape[0]
    return v.reshape(sh)


def absorber_sigma(n, nabs, dx, dtype, R0=1e-15, m=2):
    """Matched-absorber ramp (pml.py:62-80): on each absorbing edge of nc
    cells, sigma_i = smax ((nc - i - 1/2) / nc)^m, i = 0 at the wall, with
    smax = -(m + 1) ln(R0) / (2 nc dx); zero elsewhere."""
    lo, hi = _sides(nabs)
    p = torch.zeros(n, dtype=dtype)

    def ramp(nc):
        smax = -(m + 1) * math.log(R0) / (2.0 * nc * dx)
        return smax * ((nc - torch.arange(nc, dtype=dtype) - 0.5) / nc) ** m

    if lo > 0:
        p[:lo] = ramp(lo)
    if hi > 0:
        p[n - hi:] = torch.flip(ramp(hi), [0])
    return p


def _media3d(c, nx, ny, nz, dtype, sigma_d, sigma_b, poles, pole_masks, sig_bound=None):
    """stepper_3d.py:217-266.  Damping pairs of the Yee-averaged conductivity
    (E: averaged along the component's own axis; H: along the two others):
      d1 = (1 - s dt/2) / (1 + s dt/2),  d2 = dt / (1 + s dt/2);
    poles: per E component (c1, c2) scalars shaped (np, 1, 1, 1) and
    c3 = v3 * (mask averaged along the component axis), (np, *comp shape)."""
    dt = c["dt"]
    z = torch.zeros(nx, ny, nz, dtype=dtype)
    se = sigma_d.to(dtype) if sigma_d is not None else z
    sh = sigma_b.to(dtype) if sigma_b is not None else z
    if sig_bound is not None:   # stepper_3d.py:224-225: boundary sigma + material sigma
        se, sh = sig_bound + se, sig_bound + sh
    pair = lambda s: ((1 - s * dt / 2) / (1 + s * dt / 2), dt / (1 + s * dt / 2))
    c["damp"] = {"Ex": pair(avg(se, 0)), "Ey": pair(avg(se, 1)), "Ez": pair(avg(se, 2)),
                 "Hx": pair(avg(avg(sh, 1), 2)), "Hy": pair(avg(avg(sh, 0), 2)),
                 "Hz": pair(avg(avg(sh, 0), 1))}
    c["pols"] = None
    c["sigma"] = sigma_d is not None or sigma_b is not None
    if len(poles):
        rec = [pole_recurrence(p, dt) for p in poles]
        n = len(poles)
        c1 = torch.tensor([r[0] for r in rec], dtype=dtype).reshape(n, 1, 1, 1)
        c2 = torch.tensor([r[1] for r in rec], dtype=dtype).reshape(n, 1, 1, 1)
        c["pols"] = {}
        for comp, ax in (("Ex", 0), ("Ey", 1), ("Ez", 2)):
            c3 = torch.stack([rec[q][2] * avg((pole_masks[q] if pole_masks is not None
                                              else torch.ones(nx, ny, nz)).to(dtype), ax)
                              for q in range(n)])
            c["pols"][comp] = (c1, c2, c3)


def coeffs3d(nx, ny, nz, dx, eps, npml, dtype=torch.float32, sigma_d=None, sigma_b=None, poles=(),
             pole_masks=None, n_abs=0, absorber_axes=(), absorber_m=2, absorber_R0=1e-15, absorber=None):
    """Per-axis CPML profiles: H side length n-1 at offset 1/2, E side length n
    at offset 0 trimmed to [1:-1] (stepper_3d.py:182-195); dt = dx/(2 sqrt 3).
    npml / n_abs: an int for every face or per axis an int or (lo, hi) pair.
    Boundary modes (stepper_3d.py:174-215): no absorber -> CPML profiles
    (npml 0 on an axis gives b = 1, a = 0 there: psi stays zero, so npml = 0
    everywhere is a PEC box in CPML arithmetic); n_abs with no npml -> the
    absorber path (c["absorber"], no CPML) with
    sig_bound = max(sx[i], sy[j], sz[k]) of the absorber ramps; both ->
    mixed: the identity profile (b = a = 0) on absorber_axes and
    sig_bound = max of their ramps.  absorber=True forces the absorber path
    with sig_bound = 0 (a run with neither, the reference's cpml=None
    branch).  With any of sigma_d, sigma_b (cell-centred (nx, ny, nz)),
    poles or a boundary sigma the media coefficients of _media3d are
    added."""
    dt = 0.5 * dx / math.sqrt(3.0)
    per = lambda v: tuple(_sides(x) for x in v) if isinstance(v, (tuple, list)) else (_sides(v),) * 3
    npml_ax, nabs_ax = per(npml), per(n_abs)
    has_pml = any(l or h for l, h in npml_ax)
    has_abs = any(l or h for l, h in nabs_ax)
    dims = (nx, ny, nz)
    ramp = lambda ax: absorber_sigma(dims[ax], nabs_ax[ax], dx, dtype, R0=absorber_R0, m=absorber_m)
    eps = eps.to(dtype)
    c = {"dt": dt, "dx": dx, "eps": eps, "pro": None, "absorber": False,
         "eps_c": [avg(eps, 0), avg(eps, 1), avg(eps, 2)],
         "inv_eps": [1.0 / avg(eps, 0), 1.0 / avg(eps, 1), 1.0 / avg(eps, 2)],
         "damp": None, "pols": None}
    sig_bound = None
    if not absorber and (has_pml or not has_abs):
        ident = set(absorber_axes) if has_abs else set()
        pro = {}
        for ax, n in enumerate(dims):
            if ax in ident:      # stepper_3d.py:132-145: b = 0, a = 0, k = 1
                z = lambda m: torch.zeros(m, dtype=dtype)
                pro[("H", ax)] = (_along(z(n - 1), ax), _along(z(n - 1), ax))
                pro[("E", ax)] = (_along(z(n - 2), ax), _along(z(n - 2), ax))
                continue
            b, a, _ = cpml_profiles(n - 1, npml_ax[ax], 0.5, dt, dx, dtype)
            pro[("H", ax)] = (_along(b, ax), _along(a, ax))
            b, a, _ = cpml_profiles(n, npml_ax[ax], 0.0, dt, dx, dtype)
            pro[("E", ax)] = (_along(b[1:-1], ax), _along(a[1:-1], ax))
        c["pro"] = pro
        if has_abs:
            sig_bound = torch.zeros(nx, ny, nz, dtype=dtype)
            for ax in ident:
                sig_bound = torch.maximum(sig_bound, _along(ramp(ax), ax).expand(nx, ny, nz))
    else:
        c["absorber"] = True
        if has_abs:
            sig_bound = torch.maximum(torch.maximum(ramp(0)[:, None, None], ramp(1)[None, :, None]),
                                      ramp(2)[None, None, :])
        else:
            sig_bound = torch.zeros(nx, ny, nz, dtype=dtype)
    if sigma_d is not None or sigma_b is not None or len(poles) or sig_bound is not None:
        _media3d(c, nx, ny, nz, dtype, sigma_d, sigma_b, poles, pole_masks, sig_bound)
    return c


PSI = {  # name: (field differentiated, axis, side)  -- stepper_3d.py:47-58
    "psi_Ez_y": ("Ez", 1, "H"), "psi_Ey_z": ("Ey", 2, "H"),
    "psi_Ex_z": ("Ex", 2, "H"), "psi_Ez_x": ("Ez", 0, "H"),
    "psi_Ey_x": ("Ey", 0, "H"), "psi_Ex_y": ("Ex", 1, "H"),
    "psi_Hz_y": ("Hz", 1, "E"), "psi_Hy_z": ("Hy", 2, "E"),
    "psi_Hx_z": ("Hx", 2, "E"), "psi_Hz_x": ("Hz", 0, "E"),
    "psi_Hy_x": ("Hy", 0, "E"), "psi_Hx_y": ("Hx", 1, "E"),
}
COMP_SHAPE = {"Ex": (-1, 0, 0), "Ey": (0, -1, 0), "Ez": (0, 0, -1),
              "Hx": (0, -1, -1), "Hy": (-1, 0, -1), "Hz": (-1, -1, 0)}


def state3d(nx, ny, nz, dtype=torch.float32, n_poles=0):
    """Fields and psi (stepper_3d.py:269-292); with n_poles the per-component
    P_cur / P_prev stacks (n_poles, *E component shape)."""
    z = lambda sh: torch.zeros(*sh, dtype=dtype)
    st = {}
    for c, d in COMP_SHAPE.items():
        st[c] = z((nx + d[0], ny + d[1], nz + d[2]))
    for name, (f, ax, _) in PSI.items():
        sh = list(st[f].shape)
        sh[ax] -= 1
        st[name] = z(sh)
    if n_poles:
        for a in "xyz":
            st[f"P{a}_cur"] = z((n_poles,) + tuple(st["E" + a].shape))
            st[f"P{a}_prev"] = z((n_poles,) + tuple(st["E" + a].shape))
    return st


def _ce(s, c, name):
    """d/dx + psi with the psi recursion psi' = b psi + a d (stepper_3d.py:368-370)."""
    f, ax, side = PSI[name]
    d = diff(s[f], ax) / c["dx"]
    b, a = c["pro"][(side, ax)]
    p = b * s[name] + a * d
    s[name] = p
    return d + p


def _trim(t, ax):
    return t.narrow(ax, 1, t.shape[ax] - 2)


def _pad(t, axes):
    pad = [0] * 6
    for ax in axes:
        pad[2 * (2 - ax)] = pad[2 * (2 - ax) + 1] = 1
    return torch.nn.functional.pad(t, pad)


def step3d(st, c, amp=0.0, src=None, lean=False):
    """One CPML leapfrog step.  H from curl E, then E from curl H with the
    PEC tangential trim (stepper_3d.py:363-437).  E update:
      functional form  E = (E*eps_c + dt*curl) / eps_c   (d1 = 1, d2 = dt)
      lean form        E = E + (curl * inv_eps) * dt      (stepper_3d_lean.py:260)
      lean with poles  E = E + ((curl*dt + sum P(n)) - sum P(n+1)) * inv_eps,
                       P(n+1) = fma(c3, E, fma(c1, P(n), c2 P(n-1))) (torch's
                       addcmul_ rounds each multiply-add once)
                       (stepper_3d_lean.py:207-231; H takes the (1, dt) pair)
    Point source (comp, i, j, k): functional ADDs (dt*amp) * (1/eps_c) at the
    node (stepper_3d.py:475-481); lean SUBTRACTS (dt*amp) * inv_eps
    (stepper_3d_lean.py:294-296).  Plane (comp, mask): E -= (dt*amp)*mask/eps_c
    (stepper_3d.py:452-461) / E -= mask*(dt*amp)*inv_eps (lean)."""
    s = dict(st)
    dt = c["dt"]
    dm, pols = c.get("damp"), c.get("pols")
    if lean and (c.get("sigma") or c.get("absorber")):
        raise ValueError("the lean form has no conductivity and no absorber")
    pn = {}
    if pols is not None:       # stepper_3d.py:318-329: P(n+1) from the incoming E(n)
        for comp in ("Ex", "Ey", "Ez"):
            c1, c2, c3 = pols[comp]
            a = comp[1]
            if lean:   # stepper_3d_lean.py:207-216: in-place mul_ / addcmul_, which fuse each multiply-add
                pn[comp] = (c2 * st[f"P{a}_prev"]).addcmul_(c1, st[f"P{a}_cur"]).addcmul_(c3, st[comp])
            else:
                pn[comp] = c1 * st[f"P{a}_cur"] + c2 * st[f"P{a}_prev"] + c3 * st[comp]
            s[f"P{a}_cur"], s[f"P{a}_prev"] = pn[comp], st[f"P{a}_cur"]
    if c.get("absorber"):
        return _absorber_step(st, s, c, pn, amp, src)
    for comp, pa, pb in (("Hx", "psi_Ez_y", "psi_Ey_z"), ("Hy", "psi_Ex_z", "psi_Ez_x"),
                         ("Hz", "psi_Ey_x", "psi_Ex_y")):
        curl = _ce(s, c, pa) - _ce(s, c, pb)
        if dm is not None:     # stepper_3d.py:375,382,389: H = d1 H - d2 curl E
            s[comp] = dm[comp][0] * st[comp] - dm[comp][1] * curl
        else:
            s[comp] = st[comp] - dt * curl
    # E: curl terms (each psi updated on its full extent, then trimmed)
    for comp, (pa, fa), (pb, fb), axes in (
            ("Ex", ("psi_Hz_y", 2), ("psi_Hy_z", 1), (1, 2)),
            ("Ey", ("psi_Hx_z", 0), ("psi_Hz_x", 2), (0, 2)),
            ("Ez", ("psi_Hy_x", 1), ("psi_Hx_y", 0), (0, 1))):
        ta = _trim(_ce(s, c, pa), fa)
        tb = _trim(_ce(s, c, pb), fb)
        curl = _pad(ta - tb, axes)
        k = "xyz".index(comp[1])
        if lean and pols is not None:   # stepper_3d_lean.py:207-231 (_pole_step, _pole_e_lean)
            a = comp[1]
            s[comp] = st[comp] + (((curl * dt) + st[f"P{a}_cur"].sum(0)) - pn[comp].sum(0)) * c["inv_eps"][k]
        elif lean:
            s[comp] = st[comp] + curl * c["inv_eps"][k] * dt
        elif pols is not None:  # stepper_3d.py:295-301 (_e_pole)
            e = c["eps_c"][k]
            d1, d2 = dm[comp]
            s[comp] = (d1 * (st[comp] * e + st[f"P{comp[1]}_cur"].sum(0)) + d2 * curl
                       - pn[comp].sum(0)) / e
        elif dm is not None:    # stepper_3d.py:402-405: E = (d1 E eps_c + d2 curl) / eps_c
            e = c["eps_c"][k]
            d1, d2 = dm[comp]
            s[comp] = (d1 * st[comp] * e + d2 * curl) / e
        else:
            e = c["eps_c"][k]
            s[comp] = (st[comp] * e + dt * curl) / e
    return _source(s, c, amp, src, lean)


def _absorber_step(st, s, c, pn, amp, src):
    """stepper_3d.py:328-361 (no CPML): H = d1 H - d2 (dE_a - dE_b) / dx
    (Python order: (d2 * diff) / dx), E curl = pad((dH_a - dH_b) / dx) over
    the interior, then the same E updates as the CPML path."""
    dx, dm, pols = c["dx"], c["damp"], c.get("pols")
    for comp, (fa, aa), (fb, ab) in (("Hx", ("Ez", 1), ("Ey", 2)), ("Hy", ("Ex", 2), ("Ez", 0)),
                                     ("Hz", ("Ey", 0), ("Ex", 1))):
        d1, d2 = dm[comp]
        s[comp] = d1 * st[comp] - d2 * (diff(st[fa], aa) - diff(st[fb], ab)) / dx
    for comp, (fa, aa, ta), (fb, ab, tb), axes in (
            ("Ex", ("Hz", 1, 2), ("Hy", 2, 1), (1, 2)),
            ("Ey", ("Hx", 2, 0), ("Hz", 0, 2), (0, 2)),
            ("Ez", ("Hy", 0, 1), ("Hx", 1, 0), (0, 1))):
        curl = _pad((_trim(diff(s[fa], aa), ta) - _trim(diff(s[fb], ab), tb)) / dx, axes)
        k = "xyz".index(comp[1])
        e = c["eps_c"][k]
        d1, d2 = dm[comp]
        if pols is not None:
            s[comp] = (d1 * (st[comp] * e + st[f"P{comp[1]}_cur"].sum(0)) + d2 * curl - pn[comp].sum(0)) / e
        else:
            s[comp] = (d1 * st[comp] * e + d2 * curl) / e
    return _source(s, c, amp, src, False)


def _source(s, c, amp, src, lean):
    """Soft sources of step3d (stepper_3d.py:448-481, stepper_3d_lean.py:294-296)."""
    dt = c["dt"]
    if src is not None and amp != 0.0:
        comp = src[0]
        k = "xyz".index(comp[1])
        if len(src) == 4:
            i, j, kk = src[1:]
            E = s[comp].clone()
            if lean:
                E[i, j, kk] = E[i, j, kk] - dt * amp * c["inv_eps"][k][i, j, kk]
            else:
                E[i, j, kk] = E[i, j, kk] + dt * amp / c["eps_c"][k][i, j, kk]
            s[comp] = E
        else:
            mask = src[1].to(s[comp].dtype)
            if lean:
                s[comp] = s[comp] - mask * (dt * amp) * c["inv_eps"][k]
            else:
                s[comp] = s[comp] - dt * amp * mask / c["eps_c"][k]
    return s


def flux_plane_sample(st, normal, idx, freqs, dt, t_idx, ap=0, win=None):
    """monitors_3d.py:124-181: 8 phasor parts of (Ea, Ha, Eb, Hb) on an
    axis-normal plane, cropped to the common shape and aperture."""
    if normal == "z":
        Ea, Ha = st["Ex"][:, :, idx], 0.5 * (st["Hy"][:, :, idx - 1] + st["Hy"][:, :, idx])
        Eb, Hb = st["Ey"][:, :, idx], 0.5 * (st["Hx"][:, :, idx - 1] + st["Hx"][:, :, idx])
    elif normal == "x":
        Ea, Ha = st["Ey"][idx], 0.5 * (st["Hz"][idx - 1] + st["Hz"][idx])
        Eb, Hb = st["Ez"][idx], 0.5 * (st["Hy"][idx - 1] + st["Hy"][idx])
    else:
        Ea, Ha = st["Ez"][:, idx], 0.5 * (st["Hx"][:, idx - 1] + st["Hx"][:, idx])
        Eb, Hb = st["Ex"][:, idx], 0.5 * (st["Hz"][:, idx - 1] + st["Hz"][:, idx])
    ts = [Ea, Ha, Eb, Hb]
    m0 = min(t.shape[0] for t in ts)
    m1 = min(t.shape[1] for t in ts)
    l0, h0, l1, h1 = win if win is not None else (ap, m0 - ap, ap, m1 - ap)
    l0, l1, h0, h1 = max(0, l0), max(0, l1), min(m0, h0), min(m1, h1)
    Ea, Ha, Eb, Hb = (t[l0:h0, l1:h1] for t in ts)
    dtype = Ea.dtype
    w = (2 * math.pi * torch.tensor(freqs, dtype=dtype))[:, None, None]
    tE, tH = (t_idx + 1.0) * dt, (t_idx + 0.5) * dt
    parts = [torch.cos(w * tE) * Ea, -torch.sin(w * tE) * Ea,
             torch.cos(w * tH) * Ha, -torch.sin(w * tH) * Ha,
             torch.cos(w * tE) * Eb, -torch.sin(w * tE) * Eb,
             torch.cos(w * tH) * Hb, -torch.sin(w * tH) * Hb]
    return dt * torch.cat([p.unsqueeze(1) for p in parts], dim=1)


def flux_plane(acc, dx):
    """monitors_3d.py:183-189: 1/2 sum(Re a E Re a H + ... - b pair) dx^2."""
    p = acc.unbind(1)
    sa = p[0] * p[2] + p[1] * p[3]
    sb = p[4] * p[6] + p[5] * p[7]
    return 0.5 * (sa - sb).sum(tuple(range(1, sa.ndim))) * dx * dx


def dft_plane_sample(st, comp, plane, freq, dt, t_idx):
    """monitors_3d.py:81-94 (single frequency; the reference's reshape admits
    only one): dt * [cos(wt) F, -sin(wt) F] shaped (1, 2, 1, ...)."""
    F = st[comp]
    if plane is not None:
        sl = [slice(None)] * 3
        sl[plane[0]] = plane[1]
        F = F[tuple(sl)]
    off = 1.0 if comp.startswith("E") else 0.5
    t = (t_idx + off) * dt
    w = (2 * math.pi * torch.tensor([freq], dtype=F.dtype)).reshape([1, 1] + [1] * F.ndim)
    return dt * torch.stack([torch.cos(w * t) * F, -torch.sin(w * t) * F], dim=1)


def run3d(st, c, n_steps, amps=None, src=None, monitors=(), lean=False):
    """monitors_3d.py:192-210 / stepper_3d_lean.py:321-348: sequential sums.
    monitors: {"type": "flux", "normal", "idx", "freqs", "ap", "win"} or
    {"type": "dft", "comp", "plane", "freq"}."""
    acc = [None] * len(monitors)
    for t in range(n_steps):
        st = step3d(st, c, amps[t] if amps is not None else 0.0, src, lean=lean)
        for i, m in enumerate(monitors):
            if m["type"] == "flux":
                smp = flux_plane_sample(st, m["normal"], m["idx"], m["freqs"], c["dt"], t,
                                        m.get("ap", 0), m.get("win"))
            else:
                smp = dft_plane_sample(st, m["comp"], m.get("plane"), m["freq"], c["dt"], t)
            acc[i] = smp if acc[i] is None else acc[i] + smp
    return st, acc
