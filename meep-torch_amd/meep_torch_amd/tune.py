"""Autotuner for the temporally blocked 2D steppers.

For each problem signature (mode, grid, dtype, monitor layout) the first run
times every tile variant (owned tile shape x Yee steps per launch, see
mt_tiled_variants) and the two-pass per-half-step path on scratch copies of
the real state, sources and monitors, and caches the fastest for the rest of
the process.  MEEP_TORCH_AMD_VARIANT=<v> forces a variant (-1 = two-pass)."""
import os

import torch

from . import _engine as E

_CACHE = {}
MIN_CELLS = 96 * 96          # below this, launch overhead dominates: no tuning
TUNE_STEPS = 24
DEFAULT_VARIANT = 1


def variants(dtype):
    L = E.lib()
    n = L.mt_tiled_variants(E.dtype_code(dtype), None, None, None)
    steps = (E.c_int * n)()
    tx = (E.c_int * n)()
    ty = (E.c_int * n)()
    L.mt_tiled_variants(E.dtype_code(dtype), steps, tx, ty)
    return [(int(tx[v]), int(ty[v]), int(steps[v])) for v in range(n)]


def _key(mode, fields, monitors):
    ref = fields["Ez"] if mode == "tm" else fields["Hz"]
    return (mode, tuple(ref.shape), ref.dtype, ref.device.index,
            tuple(type(m).__name__ for m in monitors))


def variant_for(mode, fields, coeffs, plan, monitors, k):
    env = os.environ.get("MEEP_TORCH_AMD_VARIANT")
    if env is not None:
        return int(env)
    ref = fields["Ez"] if mode == "tm" else fields["Hz"]
    if ref.numel() < MIN_CELLS or k < 2 * TUNE_STEPS:
        return DEFAULT_VARIANT
    key = _key(mode, fields, monitors)
    if key not in _CACHE:
        _CACHE[key] = tune(mode, fields, coeffs, plan, monitors)
    return _CACHE[key]


def tune(mode, fields, coeffs, plan, monitors, steps=TUNE_STEPS):
    """Time each candidate on copies of `fields`; returns the fastest variant."""
    from .core import engine2d
    ref = fields["Ez"] if mode == "tm" else fields["Hz"]
    cands = [-1] + list(range(len(variants(ref.dtype))))
    times = {}
    for v in cands:
        work = {f: t.clone() for f, t in fields.items()}
        alt = {f: t.clone() for f, t in fields.items()} if v >= 0 else None
        engine2d.run_chunk(mode, work, coeffs, 0, 4, plan, monitors, variant=v, alt=alt)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        engine2d.run_chunk(mode, work, coeffs, 4, steps, plan, monitors, variant=v, alt=alt)
        e1.record()
        torch.cuda.synchronize()
        times[v] = e0.elapsed_time(e1)
    best = min(times, key=times.get)
    tune.last = times
    return best


tune.last = {}
