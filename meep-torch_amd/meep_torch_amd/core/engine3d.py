"""Host side of the 3D engine: one mt_yee3d_run call per run.

Field dict keys are the State3D names (reference shapes); psi arrays are the
full-grid reference tensors.  Sources follow the reference tuples
(stepper_3d.py:448-481): (comp, i, j, k) point, (comp, mask) mask,
(comp, ix, plane) plane-indexed."""
import ctypes

import torch

from .. import _engine as E
from .engine2d import stage_alloc
from .stepper_3d import COMPS, PSI_NAMES, has_media

_COMP = {"Ex": 0, "Ey": 1, "Ez": 2}

# 0: two kernels per step; 4: two kernels streaming along x; 1/2/3: one
# fused pass per step (8/4/16-row tiles)
DEFAULT_VARIANT = 0


def _source(src, fields, lean, dtype, device):
    s = E.Source3D()
    if src is None:
        return s, None
    comp = src[0]
    if comp not in _COMP:
        raise ValueError(f"3D source component must be Ex/Ey/Ez, got {comp!r}")
    s.comp = _COMP[comp]
    shape = tuple(fields[comp].shape)
    if len(src) == 4:
        i, j, k = (int(v) for v in src[1:])
        for v, n in zip((i, j, k), shape):
            if not 0 <= v < n:
                raise IndexError(f"point source {src} outside the {comp} grid {shape}")
        s.kind = 1
        s.i0, s.i1, s.j0, s.j1, s.k0, s.k1 = i, i + 1, j, j + 1, k, k + 1
        return s, None
    if len(src) == 2:
        mask = src[1].to(device=device, dtype=dtype)
        if tuple(mask.shape) != shape:
            raise ValueError(f"source mask shape {tuple(mask.shape)} != {comp} grid {shape}")
        box = E.mask_box(mask)
    elif len(src) == 3:
        ix, plane = int(src[1]), src[2].to(device=device, dtype=dtype)
        mask = torch.zeros(shape, dtype=dtype, device=device)
        # functional plane-indexed form ADDS (stepper_3d.py:462-474): negate
        # the weights so the kernel's subtraction adds exactly
        mask[ix] = plane if lean else -plane
        box = E.mask_box(mask)
    else:
        raise ValueError(f"unrecognised 3D source tuple of length {len(src)}")
    if box is None:
        return E.Source3D(), None
    (i0, j0, k0), (i1, j1, k1), w = box
    s.kind = 2
    s.i0, s.i1, s.j0, s.j1, s.k0, s.k1 = i0, i1, j0, j1, k0, k1
    s.w = w.data_ptr()
    return s, w


def _axis(c, key):
    b, a, _ = c.cpml[key]
    return E.pml_axis(b, a, c.shells[key] if c.shells else None)


def _no_shell(n):
    """Profile of an axis without CPML (absorber runs): no arrays, no shell."""
    return E.PmlAxis(None, None, n, 0, n)


def _media(d, c, fields, dtype, device, damping=True):
    """Fill the descriptor's media fields from Coeffs3D (the (d1, d2) pairs of
    c.damp and the pole scalars / c3 stacks of c.pols, stepper_3d.py:217-266;
    damping=False: the poles alone, for the lean form, stepper_3d_lean.py:
    143-165); returns the device tensors the call must keep alive."""
    keep = []

    def dev(t):
        t = t.to(device=device, dtype=dtype).contiguous()
        keep.append(t)
        return t.data_ptr()

    shapes = {"ex": fields["Ex"].shape, "ey": fields["Ey"].shape, "ez": fields["Ez"].shape,
              "hx": fields["Hx"].shape, "hy": fields["Hy"].shape, "hz": fields["Hz"].shape}
    for q, key in enumerate(("ex", "ey", "ez", "hx", "hy", "hz") if damping else ()):
        for r, t in enumerate(c.damp[key]):
            if tuple(t.shape) != tuple(shapes[key]):
                raise ValueError(f"damping pair {key} of shape {tuple(t.shape)} != {tuple(shapes[key])}")
            d.damp[2 * q + r] = dev(t)
    if c.pols is not None:
        n = int(c.pols["x"][0].numel())
        d.n_poles = n
        d.c1 = dev(c.pols["x"][0].reshape(-1))
        d.c2 = dev(c.pols["x"][1].reshape(-1))
        for a, ax in enumerate("xyz"):
            want = (n,) + tuple(fields["E" + ax].shape)
            for name in (f"P{ax}_cur", f"P{ax}_prev"):
                t = fields.get(name)
                if t is None or tuple(t.shape) != want:
                    raise ValueError(f"{n} 3D poles need {name} of shape {want} in the state")
            if tuple(c.pols[ax][2].shape) != want:
                raise ValueError(f"c3 stack for E{ax} of shape {tuple(c.pols[ax][2].shape)} != {want}")
            d.c3[a] = dev(c.pols[ax][2])
            d.P_cur[a] = fields[f"P{ax}_cur"].data_ptr()
            d.P_prev[a] = fields[f"P{ax}_prev"].data_ptr()
    return keep


def run_fields(fields, c, n_steps, source, monitors, lean=False, step0=0, tape=None, reuse=None, raw=False,
               variant=None):
    """Advance the field dict n_steps in place; returns monitor outputs.

    tape: device buffer of n_steps slots [Ex | Ey | Ez] receiving E after
    every step; reuse: raw outputs of a previous call whose accumulators keep
    summing (one sequential sum across calls); raw: return the raw
    (monitor, kind, buffer) outputs instead of the reshaped ones.  variant:
    0 two-pass, 1-3 split step: one-pass interior kernel (tile rows 6, 14,
    4) + shell cells, 4, 5 two steps per pass over the slab inside the x and
    y CPML (all of z; FP32 runs without media or tape whose monitors read x
    planes, else two-pass) + the x and y CPML cells (default:
    DEFAULT_VARIANT)."""
    from .engine2d import finish_outputs, monitor_structs
    L = E.lib()
    absorber = c.cpml is None   # stepper_3d.py:328-361: the matched-absorber path
    if absorber and lean:
        raise NotImplementedError("the lean stepper runs the CPML path only")
    ref = fields["Ex"]
    dtype, device = ref.dtype, ref.device
    code = E.dtype_code(dtype)
    E.require_device(*fields.values())
    for f, t in fields.items():
        if not t.is_contiguous():
            raise ValueError(f"state field {f} must be contiguous")
    nx, ny, nz = fields["Ey"].shape[0], fields["Ex"].shape[1], fields["Ex"].shape[2]
    d = E.Yee3D()
    d.nx, d.ny, d.nz, d.dtype, d.lean = nx, ny, nz, code, int(bool(lean))
    d.absorber = int(absorber)
    d.dt, d.dx = float(c.dt), float(c.dx)
    eps = c.eps.to(device=device, dtype=dtype).contiguous()
    d.eps = eps.data_ptr()
    for a, (kh, ke) in enumerate((("hx", "ex"), ("hy", "ey"), ("hz", "ez"))):
        n = (nx, ny, nz)[a]
        d.h[a] = _no_shell(n - 1) if absorber else _axis(c, kh)
        d.e[a] = _no_shell(n - 2) if absorber else _axis(c, ke)
    for comp in COMPS:
        setattr(d, comp, fields[comp].data_ptr())
    if not absorber:
        for q, name in enumerate(PSI_NAMES):
            if fields.get(name) is None:
                raise ValueError(f"a CPML run needs the psi array {name} in the state")
            d.psi[q] = fields[name].data_ptr()
    d.tape = tape.data_ptr() if tape is not None else None
    v = DEFAULT_VARIANT if variant is None else int(variant)
    if lean and bool(getattr(c.damp, "active", False)):
        raise NotImplementedError("3D conductivity (sigma) is not supported by the lean stepper")
    media = _media(d, c, fields, dtype, device, damping=not lean) if (has_media(c) or absorber) else []
    if media:
        v = 0
    alt = []
    if 1 <= v <= 5 and n_steps > 0:
        # ping-pong field set (psi stay in place); the two-step variants 4, 5
        # add a third set for the ring's first step
        alt = [torch.empty_like(fields[f]) for f in COMPS * (2 if v >= 4 else 1)]
        if v >= 4 and not absorber:   # scratch copies of the four z-CPML memory arrays
            alt += [torch.empty_like(fields[f]) for f in ("psi_Ey_z", "psi_Ex_z", "psi_Hy_z", "psi_Hx_z")]
        for q, t in enumerate(alt):
            d.alt[q] = t.data_ptr()
    d.variant = v if n_steps > 0 else 0
    src_t = source.src if source is not None else None
    s, wkeep = _source(src_t, fields, lean, dtype, device)
    coef = None
    if s.kind:
        vals = [c.dt * source.amplitude(t) for t in range(step0, step0 + n_steps)]
        coef = (ctypes.c_double * max(1, n_steps))(*vals)
    arr, outs, keep = monitor_structs(fields, list(monitors), n_steps, dtype, device, reuse=reuse)
    n = len(monitors)
    S, stage, nbytes = stage_alloc(L, arr, n, code, device)
    rc = L.mt_yee3d_run(ctypes.byref(d), ctypes.byref(s), coef, arr, n,
                        E.ptr(stage) if stage is not None else None, nbytes, S, step0, n_steps,
                        E.stream_handle())
    E.check(rc, "mt_yee3d_run")
    del eps, wkeep, keep, alt, media
    return outs if raw else finish_outputs(outs, fields)
