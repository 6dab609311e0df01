"""Time profiles and spatial source masks (meep_torch/core/sources.py:1-216).

Profiles are evaluated on the host, once per step, in step order (so a
stateful profile such as RandomSource draws its numbers in the reference's
order); the engine receives the per-step coefficients dt*amp as one array per
run call.  Masks keep the reference's normalisation: a segment of >1 cell
carries 1/dx, a single cell 1/dx^2 (sources.py:69-78)."""
import cmath
import math

import torch

from .grid import Grid


def cw_amplitude(src, t_idx: int) -> complex:
    """Complex CW amplitude at step t_idx (sources.py:6-18)."""
    prof = getattr(src, "profile", None)
    dt = getattr(src, "dt", None)
    if prof is None or dt is None or not hasattr(prof, "camplitude"):
        return complex(src.amplitude(t_idx))
    return prof.camplitude((t_idx + 0.5) * dt)


class GaussianPulse:
    """cos(2 pi fcen (t - t0)) exp(-(t - t0)^2 / (2 w^2)), w = 1/fwidth,
    t0 = cutoff * w (sources.py:26-36)."""

    def __init__(self, fcen: float, fwidth: float, cutoff: float = 5.0):
        self.fcen = fcen
        self.w = 1.0 / fwidth
        self.t0 = cutoff * self.w
        self.t_end = 2.0 * self.t0

    def amplitude(self, t: float) -> float:
        dt0 = t - self.t0
        return math.cos(2 * math.pi * self.fcen * dt0) * math.exp(-(dt0 ** 2) / (2 * self.w ** 2))


class CWRamp:
    """sin(2 pi f t) with a linear turn-on over t_ramp (sources.py:39-46)."""

    def __init__(self, fcen: float, t_ramp: float):
        self.fcen = fcen
        self.t_ramp = t_ramp

    def amplitude(self, t: float) -> float:
        return math.sin(2 * math.pi * self.fcen * t) * min(1.0, t / self.t_ramp)


class ContinuousSource:
    """CW carrier with a 1 - cos turn-on of slowness*width (sources.py:94-127)."""

    def __init__(self, frequency: float, width: float | None = None, slowness: float = 3.0,
                 cutoff: float = 5.0, fcen: float | None = None):
        self.fcen = float(fcen if fcen is not None else frequency)
        self.t_ramp = (slowness * width) if width is not None else max(1.0, slowness / self.fcen)
        self.cutoff = cutoff

    def _ramp(self, t: float) -> float:
        if self.t_ramp <= 0:
            return 1.0
        return 0.5 - 0.5 * math.cos(math.pi * min(1.0, t / self.t_ramp))

    def amplitude(self, t: float) -> float:
        return math.sin(2 * math.pi * self.fcen * t) * self._ramp(t)

    def camplitude(self, t: float) -> complex:
        return 1j * cmath.exp(-2j * math.pi * self.fcen * t) * self._ramp(t)


class CustomSource:
    """amplitude(t) = fn(t); a complex value contributes its real part
    (sources.py:130-143)."""

    def __init__(self, fn):
        self.fn = fn

    def amplitude(self, t: float) -> float:
        v = self.fn(t)
        return float(v.real) if isinstance(v, complex) else float(v)


class LineSource:
    """Soft current on x = ix, y in [j0, j1) (sources.py:49-91)."""

    def __init__(self, grid: Grid, ix: int, j0: int, j1: int, profile, amp_func=None, amps=None):
        if not 0 <= ix < grid.nx:
            raise ValueError(f"source ix={ix} outside grid nx={grid.nx}")
        if not 0 <= j0 < j1 <= grid.ny:
            raise ValueError(f"source segment j0={j0}, j1={j1} invalid for ny={grid.ny}")
        self.profile = profile
        self.dt = grid.dt
        mask = torch.zeros(grid.nx, grid.ny, dtype=grid.dtype)
        if amps is not None:
            mask[ix, j0:j1] = torch.as_tensor(amps, dtype=grid.dtype)
        elif amp_func is None:
            mask[ix, j0:j1] = 1.0
        else:
            for j in range(j0, j1):
                mask[ix, j] = float(amp_func((j + 0.5) * grid.dx))
        per_cell = grid.dx if (j1 - j0) > 1 else 1.0
        self.mask = (mask * (per_cell / grid.dx ** 2)).to(grid.device)

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude((t_idx + 0.5) * self.dt)


class SoftSource:
    """Jz = profile(t) * mask(x, y) for a caller-built mask (sources.py:146-159)."""

    def __init__(self, grid: Grid, mask, profile):
        self.mask = mask
        self.profile = profile
        self.dt = grid.dt

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude((t_idx + 0.5) * self.dt)


class ObliqueLineSource:
    """Phased line source (sources.py:162-193).  Its mask changes every step,
    which the engine's fixed-mask injection does not carry: run_* raises."""

    per_step_mask = True

    def __init__(self, grid: Grid, ix: int, j0: int, j1: int, profile, theta: float,
                 fcen: float | None = None):
        if not 0 <= ix < grid.nx:
            raise ValueError(f"source ix={ix} outside grid nx={grid.nx}")
        self.grid = grid
        self.ix, self.j0, self.j1 = ix, j0, j1
        self.profile = profile
        self.dt = grid.dt
        self.fcen = float(fcen if fcen is not None else getattr(profile, "fcen", 0.0))
        self.ky = 2.0 * math.pi * self.fcen * math.sin(theta)
        self.omega = 2.0 * math.pi * self.fcen
        self.mask = grid.zeros(grid.nx, grid.ny)

    def amplitude(self, t_idx: int) -> float:
        raise NotImplementedError(
            "ObliqueLineSource: per-step masks are not supported by the HIP stepper")


class RandomSource:
    """N(0, 1) amplitude per step on a fixed mask, seeded (sources.py:196-216)."""

    def __init__(self, grid: Grid, ix: int, j0: int, j1: int, amplitude: float = 1.0, seed: int = 0):
        mask = torch.zeros(grid.nx, grid.ny, dtype=grid.dtype)
        mask[ix, j0:j1] = 1.0
        self.mask = mask.to(grid.device)
        self.dt = grid.dt
        self.amp = float(amplitude)
        self.gen = torch.Generator().manual_seed(seed)

    def amplitude(self, t_idx: int) -> float:
        return self.amp * float(torch.randn(1, generator=self.gen).item())
