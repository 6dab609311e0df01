"""meep source specifications -> engine line sources
(meep_torch/mp/sources.py:100-160,484-507)."""
import math

from ..core import sources as core_sources
from .geometry import cell_index, cell_index_y, span_index_range
from .vec import Vector3


class _Delayed:
    """Profile that is zero before start_time and shifted by it afterwards."""

    def __init__(self, inner, start_time):
        self._inner = inner
        self.start_time = float(start_time)
        self.t_end = getattr(inner, "t_end", 0.0) + self.start_time

    def amplitude(self, t):
        return self._inner.amplitude(t - self.start_time) if t >= self.start_time else 0.0


class GaussianSource:
    """Gaussian pulse: width = max(width, 1/fwidth), peak at
    start_time + cutoff * width; engine profile GaussianPulse(f, 1/width, cutoff)."""

    def __init__(self, frequency=None, width=0, fwidth=float("inf"), start_time=0, cutoff=5.0,
                 is_integrated=False, wavelength=None):
        if frequency is None and wavelength is None:
            raise ValueError("GaussianSource requires either frequency or wavelength")
        if is_integrated:
            raise NotImplementedError("is_integrated=True is not supported (soft current sources only)")
        self.frequency = 1.0 / wavelength if wavelength else float(frequency)
        self.width = max(width, 1.0 / fwidth)
        self.start_time = start_time
        self.cutoff = cutoff

    def profile(self):
        if self.width == 0:
            raise ValueError("GaussianSource needs width or fwidth")
        p = core_sources.GaussianPulse(self.frequency, 1.0 / float(self.width), cutoff=self.cutoff)
        return _Delayed(p, self.start_time) if self.start_time else p


class ContinuousSource:
    def __init__(self, frequency=None, width=0, start_time=0, end_time=float("inf"), slowness=3.0,
                 wavelength=None, is_integrated=False):
        if frequency is None and wavelength is None:
            raise ValueError("ContinuousSource requires either frequency or wavelength")
        if is_integrated or start_time or not math.isinf(end_time):
            raise NotImplementedError("ContinuousSource start/end times and is_integrated are not supported")
        self.frequency = 1.0 / wavelength if wavelength else float(frequency)
        self.width = width
        self.slowness = slowness

    def profile(self):
        return core_sources.ContinuousSource(self.frequency, width=self.width or None, slowness=self.slowness)


class Source:
    def __init__(self, src, component, center=Vector3(), size=Vector3(), amplitude=1.0, amp_func=None):
        if amp_func is not None:
            raise NotImplementedError("amp_func sources are not supported")
        if complex(amplitude).imag != 0:
            raise NotImplementedError("complex source amplitudes are not supported")
        self.src, self.component = src, component
        self.center, self.size = Vector3(*center), Vector3(*size)
        self.amplitude = complex(amplitude).real


AUTOMATIC = -1      # meep: launch / decompose along the region normal, forward


class EigenModeSource(Source):
    """meep EigenModeSource (meep_torch/mp/sources.py:424-470): a port mode
    launched by the S-parameter path (EigenModeSource + add_mode_monitor ->
    ports.compute_smatrix).  The propagation axis is the axis of the region's
    smallest extent; eig_band is 1-based as in meep.  Knobs with no engine
    path raise by name."""

    def __init__(self, src, center=Vector3(), size=Vector3(0, 1, 0), component="Ez", direction=AUTOMATIC,
                 eig_band=1, amplitude=1.0, eig_kpoint=None, eig_match_freq=True, eig_parity=0, eig_vol=None,
                 eig_resolution=0, eig_tolerance=1e-12, eig_lattice_size=None, eig_lattice_center=None,
                 volume=None):
        if eig_band < 1:
            raise ValueError(f"eig_band is 1-based like meep (1 = fundamental mode), got {eig_band}")
        for name, v, ok in (("eig_kpoint", eig_kpoint, None), ("eig_vol", eig_vol, None),
                            ("eig_lattice_size", eig_lattice_size, None),
                            ("eig_lattice_center", eig_lattice_center, None), ("eig_parity", eig_parity, 0)):
            if v != ok:
                raise NotImplementedError(f"EigenModeSource {name} is not supported (modes are solved on the "
                                          "port line along its normal)")
        if eig_match_freq is not True:
            raise NotImplementedError("EigenModeSource eig_match_freq=False (fixed-k mode) is not supported")
        if volume is not None:
            center, size = volume.center, volume.size
        super().__init__(src, component, center=center, size=size, amplitude=amplitude)
        self.direction, self.eig_band = direction, int(eig_band)


def port_line(center, size, gs):
    """(axis, idx, lo, hi) of a port or mode-monitor line (meep_torch/mp/
    sources.py:510-529): normal to the region's smallest extent, spanning
    the other extent clipped to the cell."""
    c, s = center, size
    if abs(s.x) <= abs(s.y):
        y0, y1 = c.y - s.y / 2 + gs.sy / 2, c.y + s.y / 2 + gs.sy / 2
        return ("x", cell_index(c.x, gs)) + tuple(span_index_range(max(0.0, y0), min(gs.sy, y1), gs, "y"))
    x0, x1 = c.x - s.x / 2 + gs.sx / 2, c.x + s.x / 2 + gs.sx / 2
    return (("y", max(0, min(gs.ny, cell_index_y(c.y, gs))))
            + tuple(span_index_range(max(0.0, x0), min(gs.sx, x1), gs, "x")))


def source_span(spec, gs):
    """(ix, j0, j1): x-normal line (or point) of the source (meep_torch/mp/sources.py:484-501)."""
    c = spec.center
    ix = cell_index(c.x, gs)
    if spec.size.y == 0:
        jy = max(0, min(gs.ny, cell_index_y(c.y, gs)))
        return ix, jy, jy + 1
    if math.isinf(spec.size.y):
        return ix, 0, gs.ny
    y0 = c.y - spec.size.y / 2.0 + gs.sy / 2.0
    y1 = c.y + spec.size.y / 2.0 + gs.sy / 2.0
    j0, j1 = span_index_range(max(0.0, y0), min(gs.sy, y1), gs, "y")
    return ix, j0, j1
