"""Run loops (meep_torch/core/run.py:10-167) on the HIP engine.

run_tm / run_te keep the reference signature and result: (final state,
[one accumulated output per monitor]).  With checkpoint_every = K > 0 the run
is split into K-step chunks whose monitor partial sums are added in chunk
order, exactly as the reference folds them (run.py:56-63), so FP32 spectra
round the same way.  When a coefficient tensor requires grad (the design
chain eps <- rho), the run goes through the hand-written adjoint
(meep_torch_amd.core.adjoint2d) instead of autograd through ~60 ops per step.
"""
import torch

from . import engine2d
from .stepper_te import FIELDS as TE_FIELDS
from .stepper_te import TEState, _check as _check_te, fields_of as te_fields_of, step_te
from .stepper_tm import FIELDS as TM_FIELDS
from .stepper_tm import TMState, _check as _check_tm, fields_of as tm_fields_of, step_tm


class _Fixed:
    """Source with one fixed amplitude (single-step calls), optionally with
    TFSF magnetic masks and their fixed amplitude."""

    def __init__(self, amp, mask, h_amp=0.0, h_masks=None):
        self._amp = float(amp)
        self._h_amp = float(h_amp)
        self.mask = mask
        self.h_masks = {k: v for k, v in (h_masks or {}).items() if v is not None}

    def amplitude(self, t):
        return self._amp

    def h_amplitude(self, t):
        return self._h_amp


def _mode_of(step_fn, state):
    if step_fn is step_tm or isinstance(state, TMState):
        return "tm"
    if step_fn is step_te or isinstance(state, TEState):
        return "te"
    raise NotImplementedError(f"step function {step_fn!r} has no HIP engine")


def _fields(mode, state=None):
    if mode == "tm":
        return tm_fields_of(state) if state is not None else TM_FIELDS
    return te_fields_of(state) if state is not None else TE_FIELDS


def _hshapes(mode, work):
    nx, ny = (work["Ez"].shape if mode == "tm" else (work["Ey"].shape[0], work["Ex"].shape[1]))
    return engine2d.h_shapes(mode, nx, ny)


def _target_shape(mode, work):
    return work["Dz"].shape if mode == "tm" else work["Dy"].shape


def _validate(mode, state, coeffs):
    if mode == "tm":
        _check_tm(state, coeffs)
    else:
        _check_te(state, coeffs)


def _absorb(acc, monitors, outs):
    for i, (mon, out) in enumerate(zip(monitors, outs)):
        if acc[i] is None:
            acc[i] = out
        elif mon.reduce == "sum":
            acc[i] = acc[i] + out
        else:
            acc[i] = torch.cat([acc[i], out])


def _run_engine(mode, state, coeffs, n_steps, source, monitors, checkpoint_every):
    _validate(mode, state, coeffs)
    names = _fields(mode, state)
    work = {f: getattr(state, f).detach().clone().contiguous() for f in names}
    plan = engine2d.source_plan(source, _target_shape(mode, work), work[names[0]].device,
                                work[names[0]].dtype, _hshapes(mode, work))
    acc = [None] * len(monitors)
    chunk = n_steps if checkpoint_every <= 0 else checkpoint_every
    step0 = 0
    while step0 < n_steps:
        k = min(chunk, n_steps - step0)
        outs = engine2d.run_chunk(mode, work, coeffs, step0, k, plan, monitors)
        _absorb(acc, monitors, outs)
        step0 += k
    return state._replace(**work), acc


def _wants_grad(coeffs):
    return any(torch.is_tensor(t) and t.requires_grad for t in coeffs)


def _make_runner(step_fn, state_cls):
    mode = "tm" if step_fn is step_tm else "te"

    def run(state, coeffs, n_steps: int, *, source=None, monitors=(), checkpoint_every: int = 0):
        monitors = list(monitors)
        if _wants_grad(coeffs) and torch.is_grad_enabled():
            from .adjoint2d import run_with_adjoint
            return run_with_adjoint(mode, state, coeffs, n_steps, source, monitors,
                                    checkpoint_every)
        with torch.no_grad():
            return _run_engine(mode, state, coeffs, n_steps, source, monitors, checkpoint_every)

    run.__name__ = f"run_{mode}"
    run.__doc__ = (f"run_{mode}(state, coeffs, n_steps, *, source=None, monitors=(), "
                   f"checkpoint_every=0) -> (state, accs) (meep_torch/core/run.py:44-120)")
    return run


run_tm = _make_runner(step_tm, TMState)
run_te = _make_runner(step_te, TEState)


def _run_chunk(step_fn, state, coeffs, step0: int, k: int, source, monitors):
    """run.py:10-40: advance k steps from global index step0; returns
    (state, [chunk output per monitor])."""
    mode = _mode_of(step_fn, state)
    _validate(mode, state, coeffs)
    names = _fields(mode, state)
    work = {f: getattr(state, f).detach().clone().contiguous() for f in names}
    plan = engine2d.source_plan(source, _target_shape(mode, work), work[names[0]].device,
                                work[names[0]].dtype, _hshapes(mode, work))
    with torch.no_grad():
        outs = engine2d.run_chunk(mode, work, coeffs, step0, k, plan, list(monitors))
    return state._replace(**work), outs


def run_until_decayed(step_fn, state, coeffs, source, monitors, probe_fn,
                      decay_ratio: float = 1e-6, n_decay: int = 50, max_steps: int = 100000,
                      block: int = 1):
    """run.py:126-167: step until probe_fn(state) stays below decay_ratio *
    peak for n_decay consecutive steps.  The probe is a host callable, so the
    engine advances one step per call here (block > 1 checks every block steps)."""
    monitors = list(monitors)
    acc = [None] * len(monitors)
    peak, decay_run, t = 0.0, 0, 0
    while t < max_steps:
        k = min(block, max_steps - t)
        state, outs = _run_chunk(step_fn, state, coeffs, t, k, source, monitors)
        _absorb(acc, monitors, outs)
        t += k
        v = float(probe_fn(state))
        if v > peak:
            peak, decay_run = v, 0
        elif peak > 0 and v < decay_ratio * peak:
            decay_run += k
            if decay_run >= n_decay:
                break
        else:
            decay_run = 0
    return state, acc
