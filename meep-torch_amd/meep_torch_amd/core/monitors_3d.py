"""3D sources, monitors and the forward runner (meep_torch/core/monitors_3d.py).

Same constructors, result shapes and flux formula as the reference; sampling
runs inside the engine (four cropped plane terms per flux plane, one term per
DFT plane)."""

import torch

from .engine3d import run_fields
from .stepper_3d import state_names


class PlaneSource:
    """Soft area current on an axis-normal plane (monitors_3d.py:18-49)."""

    def __init__(self, nx, ny, nz, dt, comp, axis, idx, profile, lo=None, hi=None, device=None):
        shape = {"Ex": (nx - 1, ny, nz), "Ey": (nx, ny - 1, nz), "Ez": (nx, ny, nz - 1)}[comp]
        mask = torch.zeros(shape, dtype=torch.float64, device=device)
        sl = [slice(None)] * 3
        sl[axis] = idx
        mask[tuple(sl)] = 1.0
        if lo is not None and hi is not None:
            sel = [slice(None)] * 3
            for a in range(3):
                if a != axis:
                    sel[a] = slice(lo, hi)
            window = mask[tuple(sel)].clone()
            mask = torch.zeros_like(mask)
            mask[tuple(sel)] = window
        self.profile = profile
        self.dt = dt
        self.src = (comp, mask)

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude((t_idx + 0.5) * self.dt)


class PointSource3D:
    """Soft point current (monitors_3d.py:52-67)."""
    _COMP = {"Ex", "Ey", "Ez"}

    def __init__(self, dt, comp, i, j, k, profile):
        if comp not in self._COMP:
            raise ValueError(f"PointSource3D comp must be Ex/Ey/Ez, got {comp!r}")
        self.dt = dt
        self.profile = profile
        self.src = (comp, int(i), int(j), int(k))

    def amplitude(self, t_idx: int) -> float:
        return self.profile.amplitude((t_idx + 0.5) * self.dt)


class DFTFieldsMonitor3D:
    """Running DFT of one component on a plane or the full grid
    (monitors_3d.py:70-94).  Result (nfreq, 2, 1, *plane_shape); the
    reference admits one frequency only, this accepts several."""
    reduce = "sum"

    def __init__(self, dt, freqs, component="Ez", plane=None, dtype=torch.float64):
        self.component = component
        self.dt = dt
        self.freqs = torch.tensor(freqs, dtype=dtype)
        self.plane = plane

    def _field(self, f):
        F = f[self.component]
        if self.plane is not None:
            axis, idx = self.plane
            sl = [slice(None)] * 3
            sl[axis] = idx
            F = F[tuple(sl)]
        return F

    def engine_terms(self, f):
        F = self._field(f)
        half = 0 if self.component.startswith("E") else 1
        if F.dim() == 3:
            return "dft", [(F.reshape(1, -1), None, half)]
        return "dft", [(F, None, half)]

    def out_shape(self, f):
        return (self.freqs.numel(), 2, 1) + tuple(self._field(f).shape)


class FluxMonitor3D:
    """Poynting flux through an axis-normal plane (monitors_3d.py:97-189)."""
    reduce = "sum"

    def __init__(self, dt, freqs, *, normal, idx, ap=0, win=None, dtype=torch.float64):
        if normal not in ("x", "y", "z"):
            raise ValueError(f"normal must be x/y/z, got {normal!r}")
        self.normal = normal
        self.idx = idx
        self.ap = ap
        self.win = win
        self.dt = dt
        self.freqs = torch.tensor(freqs, dtype=dtype)

    def _plane(self, f):
        n, k = self.normal, self.idx
        if n == "z":
            ts = [f["Ex"][:, :, k], (f["Hy"][:, :, k - 1], f["Hy"][:, :, k]),
                  f["Ey"][:, :, k], (f["Hx"][:, :, k - 1], f["Hx"][:, :, k])]
        elif n == "x":
            ts = [f["Ey"][k], (f["Hz"][k - 1], f["Hz"][k]),
                  f["Ez"][k], (f["Hy"][k - 1], f["Hy"][k])]
        else:
            ts = [f["Ez"][:, k], (f["Hx"][:, k - 1], f["Hx"][:, k]),
                  f["Ex"][:, k], (f["Hz"][:, k - 1], f["Hz"][:, k])]
        shp = [t[0].shape if isinstance(t, tuple) else t.shape for t in ts]
        m0 = min(s[0] for s in shp)
        m1 = min(s[1] for s in shp)
        if self.win is not None:
            l0, h0, l1, h1 = self.win
        else:
            l0, h0, l1, h1 = self.ap, m0 - self.ap, self.ap, m1 - self.ap
        l0, l1 = max(0, l0), max(0, l1)
        h0, h1 = min(m0, h0), min(m1, h1)
        crop = lambda t: t[l0:h0, l1:h1]
        return [tuple(crop(u) for u in t) if isinstance(t, tuple) else crop(t) for t in ts]

    def engine_terms(self, f):
        Ea, (Ha, Ha2), Eb, (Hb, Hb2) = self._plane(f)
        return "dft", [(Ea, None, 0), (Ha, Ha2, 1), (Eb, None, 0), (Hb, Hb2, 1)]

    def out_shape(self, f):
        Ea = self._plane(f)[0]
        return (self.freqs.numel(), 8) + tuple(Ea.shape)

    def flux(self, acc: torch.Tensor, dx: float) -> torch.Tensor:
        ReaE, ImaE, ReaH, ImaH, RebE, ImbE, RebH, ImbH = acc.unbind(1)
        sa = (ReaE * ReaH + ImaE * ImaH)
        sb = (RebE * RebH + ImbE * ImbH)
        return 0.5 * (sa - sb).sum(tuple(range(1, sa.ndim))) * dx * dx


def run_3d(state, coeffs, n_steps: int, *, source=None, monitors=()):
    """monitors_3d.py:192-210 on the HIP engine (functional: the input state is
    cloned)."""
    if state.psi_Ez_y is None and coeffs.cpml is not None:
        raise ValueError("a CPML run needs a state with the CPML memory (init_3d_state(npml > 0))")
    with torch.no_grad():
        work = {f: getattr(state, f).detach().clone().contiguous() for f in state_names(state)}
        outs = run_fields(work, coeffs, n_steps, source, list(monitors), lean=False)
    return state._replace(**work), outs
