"""Checkpointed 3D adjoint on the HIP engine (meep_torch/core/adjoint3d.py:40-115).

run_3d_adjoint(c, n_steps, chunk, source, monitors, loss_fn, state0) ->
(loss value, accs), with dL/deps accumulated on c.eps (and routed into
whatever graph produced it), like the reference.  Scheme:

1. forward on the HIP stepper, saving the full state on the device at every
   chunk start; the monitor accumulators are one sequential sum over all
   steps, as the reference's per-step accumulation;
2. L = loss_fn(accs) on re-leafed accumulators; autograd gives dL/dacc;
3. for each chunk, last first: restore its start state, recompute the chunk
   while the kernel tapes E after every step, then mt_yee3d_adjoint sweeps
   the chunk backwards (monitor cotangents injected at every step) and
   accumulates dL/deps.

Functional stepper (step_3d, CPML, material conductivity and ADE poles)
only; the lean stepper raises NotImplementedError.
"""
import ctypes

import torch

from .. import _engine as E
from .engine3d import _axis, run_fields
from .stepper_3d import COMPS, PSI_NAMES, state_names

__all__ = ["run_3d_adjoint"]

_G_SHAPES = ("Ex", "Ex", "Ey", "Ey", "Ez", "Ez", "Hx", "Hx", "Hy", "Hy", "Hz", "Hz")


def _fields_of(state):
    if state.psi_Ez_y is None:
        raise NotImplementedError("3D adjoint needs the CPML state (psi arrays)")
    return {f: getattr(state, f).detach().clone().contiguous() for f in state_names(state)}


def _refuse_media_grads(c):
    """The reverse sweep differentiates with respect to eps only; a material
    tensor (sigma_d, sigma_b, pole masks / c3) that asks for a gradient
    raises instead of silently receiving none."""
    damp = c.damp
    for name, t in (("sigma_d", getattr(damp, "_sd", None)), ("sigma_b", getattr(damp, "_sb", None))):
        if torch.is_tensor(t) and t.requires_grad:
            raise NotImplementedError(f"gradient with respect to {name} is not supported in 3D")
    if c.pols is not None:
        for ax, trip in c.pols.items():
            for t in trip:
                if torch.is_tensor(t) and t.requires_grad:
                    raise NotImplementedError("gradient with respect to 3D pole masks / coefficients is not supported")


def run_3d_adjoint(c, n_steps: int, chunk: int, source, monitors, loss_fn, state0):
    from .adjoint2d import _adj_monitors
    if chunk < 1:
        raise ValueError("chunk must be >= 1")
    if c.cpml is None:
        raise NotImplementedError("the hand-written 3D adjoint transposes the CPML path; a run without CPML "
                                  "(matched-absorber path) has none")
    _refuse_media_grads(c)
    L = E.lib()
    monitors = list(monitors)
    work = _fields_of(state0)
    ref = work["Ex"]
    dtype, device = ref.dtype, ref.device
    E.require_device(ref)
    eps_src = c.eps
    eps = eps_src.detach().to(device=device, dtype=dtype).contiguous()
    cd = c._replace(eps=eps)

    # 1. forward with checkpoints
    ckpts, raw = [], None
    for t0 in range(0, n_steps, chunk):
        k = min(chunk, n_steps - t0)
        ckpts.append((t0, k, {f: t.clone() for f, t in work.items()}))
        raw = run_fields(work, cd, k, source, monitors, step0=t0, reuse=raw, raw=True)
    if raw is None:
        raw = run_fields(work, cd, 0, source, monitors, raw=True)
    from .engine2d import finish_outputs
    accs = finish_outputs(raw, work)

    # 2. loss head
    leaves = [a.detach().clone().requires_grad_(True) for a in accs]
    with torch.enable_grad():
        Lval = loss_fn(leaves)
        used = [lf for lf in leaves]
        gaccs = torch.autograd.grad(Lval, used, allow_unused=True)

    # 3. reverse sweep
    A = {f: torch.zeros_like(t) for f, t in work.items()}
    scratch = [torch.empty_like(work[s]) for s in _G_SHAPES]
    geps = [torch.empty_like(work[s]) for s in ("Ex", "Ey", "Ez")]
    eps_grad = torch.zeros_like(eps)
    adj = E.Yee3DAdj()
    for f in COMPS:
        setattr(adj, "a" + f, A[f].data_ptr())
    for q, name in enumerate(PSI_NAMES):
        adj.apsi[q] = A[name].data_ptr()
        adj.g[q] = scratch[q].data_ptr()
    for q in range(3):
        adj.geps[q] = geps[q].data_ptr()
    adj.eps_grad = eps_grad.data_ptr()
    if c.pols is not None:
        if "Px_cur" not in A:
            raise ValueError("3D poles need the P_cur / P_prev stacks in the state")
        for q, ax in enumerate("xyz"):
            adj.aP_cur[q] = A[f"P{ax}_cur"].data_ptr()
            adj.aP_prev[q] = A[f"P{ax}_prev"].data_ptr()
    mon_arr, n_used, keep = _adj_monitors(A, monitors, list(gaccs), dtype, device)
    slot = sum(work[f].numel() for f in ("Ex", "Ey", "Ez"))
    tape = torch.empty((chunk + 1) * slot, dtype=dtype, device=device)
    d = E.Yee3D()
    nx, ny, nz = work["Ey"].shape[0], work["Ex"].shape[1], work["Ex"].shape[2]
    d.nx, d.ny, d.nz, d.dtype, d.lean = nx, ny, nz, E.dtype_code(dtype), 0
    d.dt, d.dx = float(c.dt), float(c.dx)
    d.eps = eps.data_ptr()
    for a, (kh, ke) in enumerate((("hx", "ex"), ("hy", "ey"), ("hz", "ez"))):
        d.h[a] = _axis(cd, kh)
        d.e[a] = _axis(cd, ke)
    damp_keep = []
    if getattr(c.damp, "active", False) or c.pols is not None:   # conductivity pairs (ones / dt without)
        for q, key in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
            for r, t in enumerate(c.damp[key]):
                t = t.to(device=device, dtype=dtype).contiguous()
                damp_keep.append(t)
                d.damp[2 * q + r] = t.data_ptr()
    if c.pols is not None:                                       # ADE poles
        d.n_poles = int(c.pols["x"][0].numel())
        for name, t in (("c1", c.pols["x"][0].reshape(-1)), ("c2", c.pols["x"][1].reshape(-1))):
            t = t.to(device=device, dtype=dtype).contiguous()
            damp_keep.append(t)
            setattr(d, name, t.data_ptr())
        for q, ax in enumerate("xyz"):
            t = c.pols[ax][2].to(device=device, dtype=dtype).contiguous()
            damp_keep.append(t)
            d.c3[q] = t.data_ptr()
    for t0, k, st in reversed(ckpts):
        for f in work:
            work[f].copy_(st[f])
        o = 0
        for f in ("Ex", "Ey", "Ez"):
            n = work[f].numel()
            tape[o:o + n].copy_(work[f].reshape(-1))
            o += n
        run_fields(work, cd, k, source, [], step0=t0, tape=tape[slot:])
        rc = L.mt_yee3d_adjoint(ctypes.byref(d), ctypes.byref(adj), E.ptr(tape), mon_arr, n_used, t0, k,
                                E.stream_handle())
        E.check(rc, "mt_yee3d_adjoint")
    del keep
    grad = eps_grad.to(dtype=eps_src.dtype, device=eps_src.device)
    if eps_src.requires_grad:
        torch.autograd.backward(eps_src, grad)
    return float(Lval.detach()), accs
