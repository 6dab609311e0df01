"""Host side of the 2D engine: turn a meep_torch-style (state, coeffs, source,
monitors) chunk into one mt_tm2d_run / mt_te2d_run call.

The state tensors are updated in place (callers clone first to keep the
reference's functional semantics).  Monitor accumulators for the chunk are
fresh zero buffers; run.py adds chunk results exactly as the reference folds
chunk partials (meep_torch/core/run.py:56-63)."""
import ctypes

import torch

from .. import _engine as E
from .monitors import omega_of

STAGE_STEPS = 64
STAGE_BUDGET = 1 << 29  # bytes of staging per call


def h_shapes(mode, nx, ny):
    """Magnetic (TFSF) mask keys of step_tm / step_te and the field each one
    feeds: key -> (field shape, slot of mt_tm2d.hsrc / mt_te2d.hsrc)."""
    if mode == "tm":
        return {"src_hx_mask": ((nx, ny - 1), 0), "src_hy_mask": ((nx - 1, ny), 1)}
    return {"src_hz_mask": ((nx - 1, ny - 1), 0)}


def source_plan(source, field_shape, device, dtype, hshapes=None):
    """Reference source protocol (run.py:17-25): .amplitude(t), .mask, and
    for a TFSF source .h_amplitude(t) with .h_masks, a dict of magnetic
    masks keyed like step_tm / step_te's arguments (``hshapes``: the keys
    this caller accepts, from h_shapes; None refuses magnetic masks)."""
    if source is None:
        return None
    if getattr(source, "parts", None) is not None:
        return _multi_plan(source, field_shape, device, dtype)
    if getattr(source, "per_step_mask", False):
        raise NotImplementedError(f"{type(source).__name__}: per-step masks are not supported")
    h_masks = getattr(source, "h_masks", None) or {}
    if h_masks and hshapes is None:
        raise NotImplementedError("TFSF magnetic source (h_masks) is not supported on this path")
    mask = getattr(source, "mask", None)
    box = None
    if mask is not None:
        if tuple(mask.shape) != tuple(field_shape):
            raise ValueError(f"source mask shape {tuple(mask.shape)} != field shape {tuple(field_shape)}")
        box = E.mask_box(mask.to(device=device, dtype=dtype))
    hboxes = []
    for key, m in h_masks.items():
        if m is None:
            continue
        if key not in hshapes:
            raise ValueError(f"unknown magnetic source mask {key!r} (expected one of {sorted(hshapes)})")
        shape, slot = hshapes[key]
        if tuple(m.shape) != tuple(shape):
            raise ValueError(f"{key} shape {tuple(m.shape)} != field shape {tuple(shape)}")
        hb = E.mask_box(m.to(device=device, dtype=dtype))
        if hb is not None:
            hboxes.append((slot, hb))
    return {"source": source, "box": box, "h": hboxes}


def _multi_plan(source, field_shape, device, dtype):
    """Several electric sources injected as one: the reference sums
    amp_i(t) * mask_i into one mask every step and injects it with amplitude
    1 (mp/cpp/core.py:584-589).  The plan keeps every part's mask cropped to
    the union box of their nonzeros; run_chunk builds the per-step sums."""
    parts = list(source.parts)
    for p in parts:
        if getattr(p, "h_masks", None):
            raise NotImplementedError("several sources with TFSF magnetic currents are not supported")
        if getattr(p, "per_step_mask", False):
            raise NotImplementedError(f"{type(p).__name__}: per-step masks are not supported")
        if tuple(p.mask.shape) != tuple(field_shape):
            raise ValueError(f"source mask shape {tuple(p.mask.shape)} != field shape {tuple(field_shape)}")
    masks = [p.mask.detach().to(device=device, dtype=dtype) for p in parts]
    nz = torch.zeros(field_shape, dtype=torch.bool, device=device)
    for m in masks:
        nz |= m.ne(0)
    if not bool(nz.any()):
        return None
    idx = nz.nonzero()
    lo = idx.min(0).values.tolist()
    hi = (idx.max(0).values + 1).tolist()
    w = torch.stack([m[lo[0]:hi[0], lo[1]:hi[1]] for m in masks]).contiguous()
    return {"source": source, "parts": parts, "multi": w, "ubox": (lo, hi), "box": None, "h": []}


class _SummedMask:
    """Source protocol of one engine call over a per-step weight stack: the
    stack already holds sum_i amp_i(t) * mask_i, injected with amplitude 1."""

    @staticmethod
    def amplitude(t):
        return 1.0


MULTI_STACK_BYTES = 1 << 27      # per-step weight stack of one engine call


def _run_chunk_multi(mode, fields, coeffs, step0, k, plan, monitors, tape, reuse, raw, variant, alt):
    """run_chunk for a several-source plan: the k steps go to the engine in
    calls of at most MULTI_STACK_BYTES of weight stack, monitors summing on
    in one sequence across the calls.  Per step the parts' amplitudes are
    taken in source order (the reference's call order); each is rounded to
    the grid dtype and multiplied into its mask, and the products are summed
    left to right, as the reference's `amp_i * mask_i` tensor ops do."""
    parts, W = plan["parts"], plan["multi"]
    lo, hi = plan["ubox"]
    n, bi, bj = W.shape
    per = max(1, MULTI_STACK_BYTES // (bi * bj * W.element_size()))
    out, done = reuse, 0
    while True:
        m = min(per, k - done)
        amps = [[p.amplitude(t) for p in parts] for t in range(step0 + done, step0 + done + m)]
        A = torch.tensor(amps, dtype=torch.float64).reshape(m, n).to(W.dtype).to(W.device)
        stack = A[:, 0, None, None] * W[0]
        for i in range(1, n):
            stack = stack + A[:, i, None, None] * W[i]
        sub = {"source": _SummedMask(), "box": (lo, hi, stack.contiguous()), "wstep": bi * bj, "h": []}
        out = run_chunk(mode, fields, coeffs, step0 + done, m, sub, monitors,
                        tape=tape[done:done + m] if tape is not None else None, reuse=out, raw=True,
                        variant=variant, alt=alt)
        done += m
        if done >= k:
            break
    return out if raw else finish_outputs(out, fields)


def coef_array(plan, dt, step0, k):
    """dt * amp(t) for t in [step0, step0+k), in double, in step order."""
    if plan is None:
        return None
    src = plan["source"]
    vals = [dt * src.amplitude(t) for t in range(step0, step0 + k)]
    return (ctypes.c_double * max(1, k))(*vals)


def coef_arrays(plan, dt, step0, k, raw=False):
    """(dt * amplitude(t), dt * h_amplitude(t)) host arrays for the k steps
    from step0, each None when that current is absent.  Per step the electric
    amplitude is taken before the magnetic one, in the reference loop's order
    (run.py:17-23), so stateful waveforms see the same call sequence.  raw:
    the electric coefficient is amplitude(t) itself (TM under sigma_d, where
    the kernel scales it by d2, stepper_tm.py:198-200)."""
    if plan is None:
        return None, None
    src = plan["source"]
    want_e, want_h = plan["box"] is not None, bool(plan.get("h"))
    ev, hv = [], []
    for t in range(step0, step0 + k):
        a = src.amplitude(t)
        if want_e:
            ev.append(a if raw else dt * a)
        if want_h:
            hv.append(dt * src.h_amplitude(t))
    n = max(1, k)
    return ((ctypes.c_double * n)(*ev) if want_e else None, (ctypes.c_double * n)(*hv) if want_h else None)


def source_struct(plan):
    s = E.Source2D()
    if plan is None or plan["box"] is None:
        s.present = 0
        return s, None
    (i0, j0), (i1, j1), w = plan["box"]
    s.present, s.i0, s.i1, s.j0, s.j1 = 1, i0, i1, j0, j1
    s.w = E.ptr(w)
    s.wstep = int(plan.get("wstep", 0))
    return s, w


def _fill_source(s, box):
    (i0, j0), (i1, j1), w = box
    s.present, s.i0, s.i1, s.j0, s.j1 = 1, i0, i1, j0, j1
    s.w = E.ptr(w)


def monitor_structs(fields, monitors, k, dtype, device, reuse=None):
    """Engine descriptors for known monitor classes; output buffers per chunk
    (``reuse``: accumulator buffers of a previous call to keep summing into)."""
    arr = (E.Monitor * max(1, len(monitors)))()
    outs, keep = [], []
    for i, mon in enumerate(monitors):
        if not hasattr(mon, "engine_terms"):
            raise NotImplementedError(
                f"monitor {type(mon).__name__} has no engine descriptor (unknown monitor class)")
        kind, terms = mon.engine_terms(fields)
        m = arr[i]
        if len(terms) > E.MAX_TERMS:
            raise ValueError("too many terms in one monitor")
        m.n_terms = len(terms)
        for q, term in enumerate(terms):
            a, b, half = term[:3]
            sem = term[3] if len(term) > 3 else None
            pa, s0, s1, n0, n1 = E.view2d(a)
            if q == 0:
                m.n0, m.n1 = n0, n1
            elif (n0, n1) != (m.n0, m.n1):
                raise ValueError("monitor terms must share one extent")
            m.term[q].a = pa
            m.term[q].b = E.view2d(b)[0] if b is not None else None
            m.term[q].s0, m.term[q].s1 = s0, s1
            m.term[q].t_half = int(half)
            if sem is not None:
                m.term[q].field, m.term[q].line, m.term[q].fix, m.term[q].lo = sem
        if kind == "probe":
            m.kind = 1
            out = torch.zeros(k, dtype=dtype, device=device)
        else:
            m.kind = 0
            om = omega_of(mon).to(device=device, dtype=dtype)
            keep.append(om)
            m.nfreq = om.numel()
            m.omega = om.data_ptr()
            if reuse is not None:
                out = reuse[i][2]
                assert out.is_contiguous() and out.numel() == m.nfreq * 2 * m.n_terms * m.n0 * m.n1
            else:
                out = torch.zeros(m.nfreq, 2 * m.n_terms, m.n0 * m.n1, dtype=dtype, device=device)
        m.acc = out.data_ptr()
        outs.append((mon, kind, out))
    return arr, outs, keep


def finish_outputs(outs, fields):
    res = []
    for mon, kind, out in outs:
        res.append(out if kind == "probe" else out.reshape(mon.out_shape(fields)))
    return res


def stage_alloc(L, arr, n, dtype_code, device):
    if n == 0:
        return 0, None, 0
    per_step = L.mt_stage_bytes(arr, n, 1, dtype_code)
    S = STAGE_STEPS
    if per_step > 0:
        S = max(1, min(STAGE_STEPS, STAGE_BUDGET // per_step))
    nbytes = L.mt_stage_bytes(arr, n, S, dtype_code)
    buf = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
    return S, buf, nbytes


def _axis(coeffs, name):
    b = getattr(coeffs, "b_" + name)
    a = getattr(coeffs, "a_" + name)
    return E.pml_axis(b, a, coeffs.shells[name] if coeffs.shells else None)


def _tm_media(d, c, fields, dtype, device):
    """Fill the media fields of an mt_tm2d from TMCoeffs (electric poles with
    the state's pole stacks, chi2 / chi3, sigma_d, sigma_b / mu); returns the
    device tensors the descriptor points to."""
    keep = []

    def put(name, t):
        t = t.to(device=device, dtype=dtype).contiguous()
        keep.append(t)
        setattr(d, name, t.data_ptr())

    if c.mc1x is not None:   # magnetic Lorentz poles (coeffs.py:177-196): B, the M stacks and their scalars
        nm = int(c.mc1x.shape[0])
        if "Bx" not in fields or int(fields["Mx_cur"].shape[0]) != nm:
            raise ValueError(f"{nm} magnetic poles need Bx, By and M stacks of {nm} in the state")
        d.n_mpoles = nm
        for name in ("mc1x", "mc2x", "mc3x", "mc1y", "mc2y", "mc3y"):
            put(name, getattr(c, name))
        for name in ("Bx", "By", "Mx_cur", "Mx_prev", "My_cur", "My_prev"):
            setattr(d, name, fields[name].data_ptr())
    d.y_mirror, d.x_mirror = int(c.y_mirror_lo), int(c.x_mirror_lo)   # symmetry cuts
    d.y_rot2, d.rot4 = int(c.y_rot2_lo), int(c.rot4_lo)
    if "Hx0" in fields:
        d.Hx0 = fields["Hx0"].data_ptr()
    if "Hy0" in fields:
        d.Hy0 = fields["Hy0"].data_ptr()
    n = int(c.c1.shape[0])
    if n:
        if "P_cur" not in fields or int(fields["P_cur"].shape[0]) != n:
            raise ValueError(f"{n} poles need P_cur / P_prev stacks of {n} in the state")
        d.n_poles = n
        for name in ("c1", "c2", "c3"):
            put(name, getattr(c, name))
        d.P_cur, d.P_prev = fields["P_cur"].data_ptr(), fields["P_prev"].data_ptr()
    if c.has_kerr or c.has_chi2:
        d.kerr = 1
        put("chi2", c.chi2 if c.chi2 is not None else torch.zeros_like(c.eps_inf))
        put("chi3", c.chi3)
    if c.d1 is not None:
        put("d1", c.d1)
        put("d2", c.d2)
    if c.h1x is not None:
        for name in ("h1x", "h2x", "h1y", "h2y"):
            put(name, getattr(c, name))
    return keep


def _te_media(d, c, fields, dtype, device):
    """Fill the media fields of an mt_te2d from TECoeffs (per-axis electric
    poles with the state's stacks, sigma_d, sigma_b / mu)."""
    keep = []

    def put(name, t):
        t = t.to(device=device, dtype=dtype).contiguous()
        keep.append(t)
        setattr(d, name, t.data_ptr())

    d.y_mirror, d.x_mirror = int(c.y_mirror_lo), int(c.x_mirror_lo)   # mirror cuts
    if c.has_kerr or c.has_chi2:
        d.kerr = 1
        for name, t in (("chi2x", c.chi2_x), ("chi3x", c.chi3_x), ("chi2y", c.chi2_y), ("chi3y", c.chi3_y)):
            put(name, t)
    if c.inv_xx_x is not None:   # tensor eps (coeffs.py:390-400)
        for name in ("inv_xx_x", "inv_xy_x", "inv_yy_y", "inv_xy_y"):
            put(name, getattr(c, name))
    if c.c3xy is not None:       # oriented poles (coeffs.py:425-455)
        put("c3xy", c.c3xy)
        put("c3yx", c.c3yx)
    if d.kerr or c.inv_xx_x is not None or c.c3xy is not None:
        kx, ky = torch.empty_like(fields["Ex"]), torch.empty_like(fields["Ey"])
        keep += [kx, ky]
        d.kx, d.ky = kx.data_ptr(), ky.data_ptr()
    for ax, n in (("x", int(c.c1x.shape[0])), ("y", int(c.c1y.shape[0]))):
        if not n:
            continue
        cur, prev = f"P{ax}_cur", f"P{ax}_prev"
        if cur not in fields or int(fields[cur].shape[0]) != n:
            raise ValueError(f"{n} E{ax} poles need {cur} / {prev} stacks of {n} in the state")
        setattr(d, f"np{ax}", n)
        for name in (f"c1{ax}", f"c2{ax}", f"c3{ax}"):
            put(name, getattr(c, name))
        setattr(d, cur, fields[cur].data_ptr())
        setattr(d, prev, fields[prev].data_ptr())
    if c.d1x is not None:
        for name in ("d1x", "d2x", "d1y", "d2y"):
            put(name, getattr(c, name))
    if c.hz1 is not None:
        put("hz1", c.hz1)
        put("hz2", c.hz2)
    return keep


def tile_counts(mode, nx, ny, coeffs, dtype, variant):
    """(interior tiles, ring tiles) of the temporally blocked launch that a
    run of this grid and CPML makes with tile `variant` (host-only query)."""
    L = E.lib()
    d = E.TM2D() if mode == "tm" else E.TE2D()
    d.nx, d.ny, d.dtype, d.variant = nx, ny, E.dtype_code(dtype), int(variant)
    names = ("hy", "hx", "ex", "ey") if mode == "tm" else ("hx", "hy", "ey", "ex")
    for n in names:
        setattr(d, n, _axis(coeffs, n))
    if mode == "tm":
        d.y_pmc = int(bool(coeffs.y_pmc))
    n_in, n_ring = ctypes.c_int(), ctypes.c_int()
    fn = L.mt_tm2d_tiles if mode == "tm" else L.mt_te2d_tiles
    E.check(fn(ctypes.byref(d), ctypes.byref(n_in), ctypes.byref(n_ring)), f"mt_{mode}2d_tiles")
    return n_in.value, n_ring.value


TM_ORDER = ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy")
TE_ORDER = ("Hz", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx")


def run_chunk(mode, fields, coeffs, step0, k, plan, monitors, tape=None, reuse=None,
              raw=False, variant=None, alt=None, static_stage=False):
    """Advance ``fields`` (dict of device tensors, reference shapes) k steps.
    tape: optional tensor receiving E after every step: (k, nx, ny) Ez for TM,
    (k, (nx-1)*ny + nx*(ny-1)) rows [Ex | Ey] for TE.
    reuse: raw outputs of a previous call whose accumulators keep summing.
    raw: return the raw (monitor, kind, buffer) triples instead of shaped outputs."""
    if plan is not None and plan.get("multi") is not None:
        return _run_chunk_multi(mode, fields, coeffs, step0, k, plan, monitors, tape, reuse, raw, variant, alt)
    L = E.lib()
    ref = fields["Ez"] if mode == "tm" else fields["Hz"]
    dtype, device = ref.dtype, ref.device
    code = E.dtype_code(dtype)
    E.require_device(*fields.values())
    if mode == "tm":
        nx, ny = fields["Ez"].shape
        d = E.TM2D()
        d.nx, d.ny, d.dtype, d.y_pmc = nx, ny, code, int(bool(coeffs.y_pmc))
        d.dt, d.dx = float(coeffs.dt), float(coeffs.dx)
        eps = coeffs.eps_inf.to(device=device, dtype=dtype).contiguous()
        d.eps = eps.data_ptr()
        d.hy, d.hx, d.ex, d.ey = (_axis(coeffs, n) for n in ("hy", "hx", "ex", "ey"))
        for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
            setattr(d, f, fields[f].data_ptr())
        if tape is not None:
            assert tape.is_contiguous() and tape.shape == (k, nx, ny) and tape.dtype == dtype
            d.tape = tape.data_ptr()
        keep_c = [eps] + _tm_media(d, coeffs, fields, dtype, device)
    else:
        nx, ny = fields["Ey"].shape[0], fields["Ex"].shape[1]
        d = E.TE2D()
        d.nx, d.ny, d.dtype = nx, ny, code
        d.dt, d.dx = float(coeffs.dt), float(coeffs.dx)
        ex_ = coeffs.eps_x.to(device=device, dtype=dtype).contiguous()
        ey_ = coeffs.eps_y.to(device=device, dtype=dtype).contiguous()
        d.eps_x, d.eps_y = ex_.data_ptr(), ey_.data_ptr()
        d.hx, d.hy, d.ey, d.ex = (_axis(coeffs, n) for n in ("hx", "hy", "ey", "ex"))
        for f in ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"):
            setattr(d, f, fields[f].data_ptr())
        if tape is not None:
            slot = (nx - 1) * ny + nx * (ny - 1)
            assert tape.is_contiguous() and tape.shape == (k, slot) and tape.dtype == dtype
            d.tape = tape.data_ptr()
            variant = -1  # the tape is written by the two-pass path only
        keep_c = [ex_, ey_] + _te_media(d, coeffs, fields, dtype, device)
    for f, t in fields.items():
        if not t.is_contiguous():
            raise ValueError(f"state field {f} must be contiguous")
    media = bool(d.n_poles or d.kerr or d.d1 or d.h1x or d.n_mpoles or d.y_mirror or d.x_mirror or d.y_rot2
                 or d.rot4) if mode == "tm" else \
        bool(d.npx or d.npy or d.d1x or d.hz1 or d.kerr or d.inv_xx_x or d.c3xy or d.y_mirror or d.x_mirror)
    if media or (mode == "te" and plan is not None and plan.get("h")):
        variant = -1  # media and TE magnetic currents are handled by the two-pass kernels
    if variant is None:
        from .. import tune
        variant = tune.variant_for(mode, fields, coeffs, plan, monitors, k)
    d.variant = int(variant)
    if variant >= 0:
        order = TM_ORDER if mode == "tm" else TE_ORDER
        if alt is None:
            alt = {f: fields[f].clone() for f in order}
        for q, f in enumerate(order):
            d.alt[q] = alt[f].data_ptr()
    src, wkeep = source_struct(plan)
    raw_amp = (coeffs.d1 is not None) if mode == "tm" else (coeffs.d1x is not None)
    coef, hcoef = coef_arrays(plan, coeffs.dt, step0, k, raw=raw_amp)
    hkeep = []
    if plan is not None and plan.get("h"):
        for slot, hb in plan["h"]:
            _fill_source(d.hsrc[slot] if mode == "tm" else d.hsrc, hb)
            hkeep.append(hb[2])
        d.hcoef = hcoef
    arr, outs, keep = monitor_structs(fields, list(monitors), k, dtype, device, reuse)
    n = len(monitors)
    if static_stage:
        # staging from the library's static device array (mt_tm2d_run_static)
        if mode != "tm":
            raise NotImplementedError("static staging is provided for TM runs")
        rc = L.mt_tm2d_run_static(ctypes.byref(d), ctypes.byref(src), coef, arr, n, step0, k, E.stream_handle())
        E.check(rc, "mt_tm2d_run_static")
        del keep_c, wkeep, keep, hkeep, hcoef
        return outs if raw else finish_outputs(outs, fields)
    S, stage, nbytes = stage_alloc(L, arr, n, code, device)
    fn = L.mt_tm2d_run if mode == "tm" else L.mt_te2d_run
    rc = fn(ctypes.byref(d), ctypes.byref(src), coef, arr, n,
            E.ptr(stage) if stage is not None else None, nbytes, S, step0, k, E.stream_handle())
    E.check(rc, f"mt_{mode}2d_run")
    del keep_c, wkeep, keep, hkeep, hcoef
    return outs if raw else finish_outputs(outs, fields)
