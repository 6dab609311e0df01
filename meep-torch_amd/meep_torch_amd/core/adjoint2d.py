"""Hand-written time-reversed adjoint for run_tm / run_te behind torch.autograd.Function.

Replaces the reference's autograd through torch.utils.checkpoint chunks
(meep_torch/core/run.py:63-118).  Forward: the HIP stepper runs the whole
interval, saving the full TM state on the device every K steps (K =
checkpoint_every, or 200).  Backward: for each chunk, last first, the chunk is
recomputed from its checkpoint while a tape records Ez after every step, then
mt_tm2d_adjoint sweeps the chunk backwards, injecting the monitor cotangents
dL/dacc and accumulating dL/deps = sum_t -(dL/dEz_t) Ez_t / eps.  The gradient
reaches whatever graph produced eps_inf (filters, projection, paint) through
autograd as usual.

Monitor accumulators are summed exactly as the forward-only run sums them (one
sequential sum per reference chunk), so forward values are identical with or
without gradients.  Supported: eps_inf (TM) or eps_x / eps_y (TE) as the only
grad-requiring coefficients; every monitor class of the 2D engine.  TE: the
tape holds [Ex | Ey] after every step and mt_te2d_adjoint accumulates
dL/deps_x and dL/deps_y, which autograd carries back through build_te_coeffs'
face averaging to eps.  Gradients with respect to the
initial state are returned for the fields (psi: inside the CPML shells)."""
import ctypes

import torch

from .. import _engine as E
from . import engine2d
from .stepper_te import fields_of as te_fields_of
from .stepper_tm import fields_of as tm_fields_of

DEFAULT_K = 200


def _adj_monitors(adj_fields, monitors, gaccs, dtype, device):
    """Monitor descriptors whose term views point into the adjoint fields and
    whose acc holds dL/dacc (skips monitors the loss does not reach)."""
    used = [(m, g) for m, g in zip(monitors, gaccs) if g is not None]
    arr = (E.Monitor * max(1, len(used)))()
    keep = []
    for i, (mon, g) in enumerate(used):
        kind, terms = mon.engine_terms(adj_fields)
        m = arr[i]
        m.n_terms = len(terms)
        for q, term in enumerate(terms):
            a, b, half = term[:3]
            pa, s0, s1, n0, n1 = E.view2d(a)
            if q == 0:
                m.n0, m.n1 = n0, n1
            m.term[q].a = pa
            m.term[q].b = E.view2d(b)[0] if b is not None else None
            m.term[q].s0, m.term[q].s1 = s0, s1
            m.term[q].t_half = int(half)
        gc = g.detach().to(device=device, dtype=dtype).contiguous()
        keep.append(gc)
        if kind == "probe":
            m.kind = 1
        else:
            m.kind = 0
            from .monitors import omega_of
            om = omega_of(mon).to(device=device, dtype=dtype).contiguous()
            keep.append(om)
            m.nfreq = om.numel()
            m.omega = om.data_ptr()
        m.acc = gc.data_ptr()
    return arr, len(used), keep


def _forward_chunks(mode, names, work, cdet, cfg):
    """Run the whole interval in place on ``work``, keeping a device copy of
    the state at the start of every chunk; returns (accumulators, checkpoints)."""
    plan, monitors, n_steps, ck = cfg["plan"], cfg["monitors"], cfg["n_steps"], cfg["ck"]
    K = ck if ck > 0 else DEFAULT_K
    acc = [None] * len(monitors)
    ckpts = []
    step0 = 0
    raw = None
    while step0 < n_steps:
        k = min(K, n_steps - step0)
        ckpts.append((step0, k, {f: work[f].clone() for f in names}))
        # one sequential sum per reference chunk: keep summing into the
        # same buffers unless checkpoint_every starts a new chunk
        reuse = raw if (ck <= 0 and raw is not None) else None
        raw = engine2d.run_chunk(mode, work, cdet, step0, k, plan, monitors, reuse=reuse, raw=True)
        if ck > 0 or step0 + k >= n_steps:
            outs = engine2d.finish_outputs(raw, work)
            for i, (mon, out) in enumerate(zip(monitors, outs)):
                if acc[i] is None:
                    acc[i] = out
                elif mon.reduce == "sum":
                    acc[i] = acc[i] + out
                else:
                    acc[i] = torch.cat([acc[i], out])
        elif any(kind == "probe" for _, kind, _ in raw):
            raise NotImplementedError("PointProbe with gradients needs checkpoint_every > 0")
        step0 += k
    return acc, ckpts


def _chunk_cotangents(monitors, gacc, step0, k):
    """Probe cotangents are per step: slice them per chunk."""
    out = []
    for mon, g in zip(monitors, gacc):
        if g is not None and mon.reduce != "sum":
            g = g[step0:step0 + k]
        out.append(g)
    return out


class _TMRun(torch.autograd.Function):
    """Differentiable in eps_inf and, when given, the chi2 / chi3 maps and the
    poles' c3 (through which pole masks enter, coeffs.py:184-190); also
    through magnetic poles (the B / M recurrence is transposed)."""

    @staticmethod
    def forward(ctx, eps, chi2, chi3, c3, cfg, *state_in):
        names = cfg["names"]
        rep = {"eps_inf": eps.detach()}
        for name, t in (("chi2", chi2), ("chi3", chi3), ("c3", c3)):
            if t is not None:
                rep[name] = t.detach()
        cdet = cfg["coeffs"]._replace(**rep)
        work = {f: t.detach().clone().contiguous() for f, t in zip(names, state_in)}
        acc, ckpts = _forward_chunks("tm", names, work, cdet, cfg)
        ctx.cfg = cfg
        ctx.cdet = cdet
        ctx.ckpts = ckpts
        ctx.save_for_backward(eps)
        ctx.n_state = len(state_in)
        ctx.extra = (chi2 is not None, chi3 is not None, c3 is not None)
        return (*[work[f] for f in names], *acc)

    @staticmethod
    def backward(ctx, *grads):
        (eps,) = ctx.saved_tensors
        cfg, cdet = ctx.cfg, ctx.cdet
        monitors, plan, names = cfg["monitors"], cfg["plan"], cfg["names"]
        nf = len(names)
        gstate, gacc = grads[:nf], list(grads[nf:])
        ref = ctx.ckpts[0][2]["Ez"]
        dtype, device = ref.dtype, ref.device
        nx, ny = ref.shape
        adj = {f: (g.detach().to(dtype).clone().contiguous() if g is not None
                   else torch.zeros_like(ctx.ckpts[0][2][f])) for f, g in zip(names, gstate)}
        scratch = {"gHy": torch.empty(nx, ny, dtype=dtype, device=device),
                   "gHx": torch.empty(nx, ny, dtype=dtype, device=device),
                   "gEx": torch.empty(nx - 1, ny, dtype=dtype, device=device),
                   "gEy": torch.empty(nx, ny - 1, dtype=dtype, device=device)}
        eps_grad = torch.zeros(nx, ny, dtype=dtype, device=device)
        L = E.lib()
        a = E.TM2DAdj()
        for f in names:
            setattr(a, "a" + f, adj[f].data_ptr())
        for f, t in scratch.items():
            setattr(a, f, t.data_ptr())
        a.eps_grad = eps_grad.data_ptr()
        d, d_keep = _descriptor(cdet, adj, dtype, device)
        kerr = bool(cdet.has_kerr or cdet.has_chi2)
        want2, want3, want_c3 = (ctx.extra[q] and ctx.needs_input_grad[q + 1] for q in range(3))
        extra = {}
        if want2:
            extra["chi2"] = torch.zeros(nx, ny, dtype=dtype, device=device)
            if kerr:
                a.chi2_grad = extra["chi2"].data_ptr()
        if want3:
            extra["chi3"] = torch.zeros(nx, ny, dtype=dtype, device=device)
            if kerr:
                a.chi3_grad = extra["chi3"].data_ptr()
        if want_c3:
            extra["c3"] = torch.zeros(cdet.c3.shape, dtype=dtype, device=device)
            if cdet.c3.shape[0]:
                a.c3_grad = extra["c3"].data_ptr()
        for step0, k, st in reversed(ctx.ckpts):
            work = {f: t.clone() for f, t in st.items()}
            a.ez_init = st["Ez"].data_ptr()
            tape = torch.empty(k, nx, ny, dtype=dtype, device=device)
            engine2d.run_chunk("tm", work, cdet, step0, k, plan, [], tape=tape)
            g_chunk = _chunk_cotangents(monitors, gacc, step0, k)
            arr, n_used, keep = _adj_monitors(adj, monitors, g_chunk, dtype, device)
            rc = L.mt_tm2d_adjoint(ctypes.byref(d), ctypes.byref(a), E.ptr(tape), arr, n_used,
                                   step0, k, E.stream_handle())
            E.check(rc, "mt_tm2d_adjoint")
            del keep, tape
        del d_keep
        g_state = [adj[f] for f in names][:ctx.n_state]
        return (eps_grad.to(eps.dtype), extra.get("chi2"), extra.get("chi3"), extra.get("c3"), None, *g_state)


def _descriptor(c, fields, dtype, device):
    d = E.TM2D()
    nx, ny = fields["Ez"].shape
    d.nx, d.ny, d.dtype, d.y_pmc = nx, ny, E.dtype_code(dtype), int(bool(c.y_pmc))
    d.y_mirror, d.x_mirror = int(c.y_mirror_lo), int(c.x_mirror_lo)   # even cuts (stepper_tm.py:134-155)
    d.dt, d.dx = float(c.dt), float(c.dx)
    eps = c.eps_inf.to(device=device, dtype=dtype).contiguous()
    d.eps = eps.data_ptr()
    d.hy, d.hx, d.ex, d.ey = (engine2d._axis(c, n) for n in ("hy", "hx", "ex", "ey"))
    keep = [eps]
    media = []
    if c.d1 is not None or c.h1x is not None:     # conductivity (sigma_d, sigma_b, mu)
        media += ["d1", "d2", "h1x", "h2x", "h1y", "h2y"]
    if c.c1.shape[0]:                              # ADE poles
        d.n_poles = int(c.c1.shape[0])
        media += ["c1", "c2", "c3"]
    if c.mc1x is not None:                         # magnetic Lorentz poles
        d.n_mpoles = int(c.mc1x.shape[0])
        media += ["mc1x", "mc2x", "mc3x", "mc1y", "mc2y", "mc3y"]
    for name in media:
        t = getattr(c, name).to(device=device, dtype=dtype).contiguous()
        keep.append(t)
        setattr(d, name, t.data_ptr())
    if c.has_kerr or c.has_chi2:                   # chi2 / chi3 fixed point (the tape holds Dz - sum P)
        d.kerr = 1
        for name, t in (("chi2", c.chi2 if c.chi2 is not None else torch.zeros_like(c.eps_inf)),
                        ("chi3", c.chi3)):
            t = t.to(device=device, dtype=dtype).contiguous()
            keep.append(t)
            setattr(d, name, t.data_ptr())
    return d, keep


_INV = ("inv_xx_x", "inv_xy_x", "inv_yy_y", "inv_xy_y")


class _TERun(torch.autograd.Function):
    """TE counterpart of _TMRun: differentiable in eps_x and eps_y (the
    face-averaged permittivities of TECoeffs, stepper_te.py:160-168), also
    through the chi2/chi3 fixed point (stepper_te.py:169-175)."""

    @staticmethod
    def forward(ctx, eps_x, eps_y, ixx, ixyx, iyy, ixyy, cfg, *state_in):
        rep = {"eps_x": eps_x.detach(), "eps_y": eps_y.detach()}
        inv = None if ixx is None else (ixx, ixyx, iyy, ixyy)
        if inv is not None:   # tensor eps: the four inverse maps
            rep.update(zip(_INV, (t.detach() for t in inv)))
        cdet = cfg["coeffs"]._replace(**rep)
        names = cfg["names"]
        work = {f: t.detach().clone().contiguous() for f, t in zip(names, state_in)}
        acc, ckpts = _forward_chunks("te", names, work, cdet, cfg)
        ctx.cfg = cfg
        ctx.cdet = cdet
        ctx.ckpts = ckpts
        ctx.save_for_backward(eps_x, eps_y)
        ctx.n_state = len(state_in)
        ctx.tensor = inv is not None
        return (*[work[f] for f in names], *acc)

    @staticmethod
    def backward(ctx, *grads):
        eps_x, eps_y = ctx.saved_tensors
        cfg, cdet = ctx.cfg, ctx.cdet
        monitors, plan = cfg["monitors"], cfg["plan"]
        names = cfg["names"]
        nf = len(names)
        gstate, gacc = grads[:nf], list(grads[nf:])
        st0 = ctx.ckpts[0][2]
        dtype, device = st0["Hz"].dtype, st0["Hz"].device
        nx, ny = st0["Ey"].shape[0], st0["Ex"].shape[1]
        adj = {f: (g.detach().to(dtype).clone().contiguous() if g is not None
                   else torch.zeros_like(st0[f])) for f, g in zip(names, gstate)}
        scratch = {"gEy": torch.empty(nx - 1, ny - 1, dtype=dtype, device=device),
                   "gEx": torch.empty(nx - 1, ny - 1, dtype=dtype, device=device),
                   "gDx": torch.empty(nx - 1, ny, dtype=dtype, device=device),
                   "gDy": torch.empty(nx, ny - 1, dtype=dtype, device=device)}
        gx = torch.zeros(nx - 1, ny, dtype=dtype, device=device)
        gy = torch.zeros(nx, ny - 1, dtype=dtype, device=device)
        L = E.lib()
        a = E.TE2DAdj()
        for f in names:
            setattr(a, "a" + f, adj[f].data_ptr())
        for f, t in scratch.items():
            setattr(a, f, t.data_ptr())
        a.epsx_grad, a.epsy_grad = gx.data_ptr(), gy.data_ptr()
        d, d_keep = _te_descriptor(cdet, nx, ny, dtype, device)
        if d.kerr or ctx.tensor or cdet.c3xy is not None:   # workspace of the fixed point's / tensor / cross transposes
            kws = torch.empty(6 * ((nx - 1) * ny + nx * (ny - 1)), dtype=dtype, device=device)
            d_keep.append(kws)
            a.kerr_ws = kws.data_ptr()
        ginv = None
        if ctx.tensor:
            ginv = [torch.zeros(nx - 1, ny, dtype=dtype, device=device), torch.zeros(nx - 1, ny, dtype=dtype, device=device),
                    torch.zeros(nx, ny - 1, dtype=dtype, device=device), torch.zeros(nx, ny - 1, dtype=dtype, device=device)]
            a.ixx_grad, a.ixy_x_grad, a.iyy_grad, a.ixy_y_grad = (t.data_ptr() for t in ginv)
        slot = (nx - 1) * ny + nx * (ny - 1)
        for step0, k, st in reversed(ctx.ckpts):
            work = {f: t.clone() for f, t in st.items()}
            tape = torch.empty(k, slot, dtype=dtype, device=device)
            engine2d.run_chunk("te", work, cdet, step0, k, plan, [], tape=tape)
            g_chunk = _chunk_cotangents(monitors, gacc, step0, k)
            arr, n_used, keep = _adj_monitors(adj, monitors, g_chunk, dtype, device)
            rc = L.mt_te2d_adjoint(ctypes.byref(d), ctypes.byref(a), E.ptr(tape), arr, n_used,
                                   step0, k, E.stream_handle())
            E.check(rc, "mt_te2d_adjoint")
            del keep, tape
        del d_keep
        g_state = [adj[f] for f in names][:ctx.n_state]
        gi = tuple(ginv) if ginv is not None else (None,) * 4
        return (gx.to(eps_x.dtype), gy.to(eps_y.dtype), *gi, None, *g_state)


def _te_descriptor(c, nx, ny, dtype, device):
    d = E.TE2D()
    d.nx, d.ny, d.dtype = nx, ny, E.dtype_code(dtype)
    d.dt, d.dx = float(c.dt), float(c.dx)
    ex_ = c.eps_x.to(device=device, dtype=dtype).contiguous()
    ey_ = c.eps_y.to(device=device, dtype=dtype).contiguous()
    d.eps_x, d.eps_y = ex_.data_ptr(), ey_.data_ptr()
    d.hx, d.hy, d.ey, d.ex = (engine2d._axis(c, n) for n in ("hx", "hy", "ey", "ex"))
    d.y_mirror, d.x_mirror = int(c.y_mirror_lo), int(c.x_mirror_lo)   # mirror cuts (stepper_te.py:87-97)
    keep = [ex_, ey_]
    media = []
    if c.d1x is not None or c.hz1 is not None:    # conductivity (sigma_d, sigma_b, mu)
        media += ["d1x", "d2x", "d1y", "d2y", "hz1", "hz2"]
    if c.c1x.shape[0]:                             # ADE poles on Ex
        d.npx = int(c.c1x.shape[0])
        media += ["c1x", "c2x", "c3x"]
    if c.c1y.shape[0]:                             # ADE poles on Ey
        d.npy = int(c.c1y.shape[0])
        media += ["c1y", "c2y", "c3y"]
    if c.c3xy is not None:                         # oriented poles (stepper_te.py:140-145)
        for name in ("c3xy", "c3yx"):
            t = getattr(c, name).to(device=device, dtype=dtype).contiguous()
            keep.append(t)
            setattr(d, name, t.data_ptr())
    if c.inv_xx_x is not None:                     # tensor eps (stepper_te.py:160-166)
        for name in _INV:
            t = getattr(c, name).to(device=device, dtype=dtype).contiguous()
            keep.append(t)
            setattr(d, name, t.data_ptr())
    if c.has_kerr or c.has_chi2:                   # the chi2/chi3 fixed point (stepper_te.py:169-175)
        d.kerr = 1
        for name, attr in (("chi2x", "chi2_x"), ("chi3x", "chi3_x"), ("chi2y", "chi2_y"), ("chi3y", "chi3_y")):
            t = getattr(c, attr).to(device=device, dtype=dtype).contiguous()
            keep.append(t)
            setattr(d, name, t.data_ptr())
    for name in media:
        t = getattr(c, name).to(device=device, dtype=dtype).contiguous()
        keep.append(t)
        setattr(d, name, t.data_ptr())
    return d, keep


def run_with_adjoint(mode, state, coeffs, n_steps, source, monitors, checkpoint_every):
    if mode == "tm" and (coeffs.y_mirror_lo == -1 or coeffs.x_mirror_lo == -1 or coeffs.y_rot2_lo
                         or coeffs.rot4_lo):
        raise NotImplementedError("gradients through odd-mirror or Rotate2 / Rotate4 symmetry cuts of TM runs "
                                  "are not supported")
    allowed = ("eps_inf", "chi2", "chi3", "c3") if mode == "tm" else ("eps_x", "eps_y") + _INV
    for name, t in coeffs._asdict().items():
        if torch.is_tensor(t) and t.requires_grad and name not in allowed:
            raise NotImplementedError(f"gradient with respect to coefficient '{name}' is not supported")
    from .run import _validate
    _validate(mode, state, coeffs)
    cfg = {"coeffs": coeffs, "monitors": list(monitors), "n_steps": n_steps, "ck": checkpoint_every}
    if mode == "tm":
        names, eps = tuple(tm_fields_of(state)), (coeffs.eps_inf,)
        tgt = state.Dz
        fn = _TMRun
    else:
        inv = tuple(getattr(coeffs, n) for n in _INV) if coeffs.inv_xx_x is not None else (None,) * 4
        names, eps = tuple(te_fields_of(state)), (coeffs.eps_x, coeffs.eps_y, *inv)
        tgt = state.Dy
        fn = _TERun
    E.require_device(*[t for t in eps if t is not None])
    nx, ny = (state.Ez.shape if mode == "tm" else (state.Ey.shape[0], state.Ex.shape[1]))
    cfg["names"] = names
    cfg["plan"] = engine2d.source_plan(source, tgt.shape, tgt.device, tgt.dtype, engine2d.h_shapes(mode, nx, ny))
    if mode == "tm":
        # the other differentiable coefficients, when something upstream wants their gradient
        extra = [t if (torch.is_tensor(t) and t.requires_grad) else None
                 for t in (coeffs.chi2, coeffs.chi3, coeffs.c3)]
        eps = (*eps, *extra)
    outs = fn.apply(*eps, cfg, *[getattr(state, f) for f in names])
    nf = len(names)
    new_state = state._replace(**dict(zip(names, outs[:nf])))
    return new_state, list(outs[nf:])
