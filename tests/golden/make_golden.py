"""Write the golden input/output vectors that pin the oracle's FP32/FP64 rounding.

Run ONLY in the build container, where the reference checkout exists:

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

It imports the reference package (``meep_torch`` from /root/reference) for the
functions whose floating-point trajectories no closed form pins -- the FP32/FP64
state after N steps of ``step_tm``/``step_te``/``step_3d``/``step_lean``, the
DFT accumulators of the flux/field monitors, the autograd gradient of a flux
objective, and the ``Simulation`` flux of the GOLDEN_PLAIN scene -- and stores
inputs + outputs as small ``.npz`` fixtures next to this file.  Nothing from the
reference is copied: the fixtures are data (arrays), the reference never travels.

Cases (all CPU, torch threads = 8):
  pml        cpml_profiles (b, a) at several (n, npml, offset), f32 + f64
  tm_small   2D TM 64x48, npml 8, waveguide + blob, point/line source, flux
             lines (x- and y-normal), DFT-field monitors, 400 steps, plus a
             chunked (checkpoint_every=150) rerun, f32 + f64
  te_small   2D TE 64x48 analogue with FluxMonitorTE / FluxMonitorTEY
  c1         C1 full size: 256x256 TM waveguide, 16-cell PML, 2000 steps,
             50-frequency flux line; also times the reference CPU run
  d3_small   3D 40x36x32, npml 6, point + plane sources, flux + DFT planes,
             step_3d (f32, f64) and step_lean (f32)
  grad_tm    autograd dJ/deps through run_tm(checkpoint_every=200), 48x48,
             1400 steps, flux objective, f64 + f32
  grad_3d    run_3d_adjoint eps gradient (30x24x20, 90 steps, chunk 16), f64
  filters    conic/gaussian filter, tanh projection, DesignRegion.paint, f64
  golden_plain  mp.Simulation GOLDEN_PLAIN cell (100x60 TM guide), f32 + f64
  media_tm   2D TM 48x40, npml 6, 240 steps, f32 + f64: (all) Lorentz + Drude
             poles with masks, chi2 and chi3 regions, sigma_d, sigma_b and mu,
             line source; (nl) chi2 + chi3 only under a strong drive
  mag_tm     2D TM 48x40, npml 6, 240 steps, f32 + f64, magnetic Lorentz poles
             (m_poles with cell masks): (mag) with two electric poles and
             chi3 under a line source, flux line; (magh) alone under a line
             source plus a TFSF-style Hy current (h_masks / h_amplitude),
             which the B phase injects into B
  media_te   2D TE analogue: shared and per-axis poles, chi2 / chi3, sigma_d,
             sigma_b, mu; (nl) chi2 + chi3 only
  tensor_te  2D TE 48x40, npml 6, 240 steps, f32 + f64, anisotropic media: (ten)
             an in-plane tensor eps (exx, exy, eyy) with an oriented pole
             (fn, gamma, (sxx, sxy, syy)) and a shared scalar pole; (tenk)
             the tensor eps under chi3; line source, flux line
  fold_2d    symmetry-folded 2D runs, 36x36, npml 6 on the far sides, 200
             steps, f32 + f64: TM with an even / odd y mirror, even / odd x
             mirror, both mirrors, Rotate2 (+-1) and Rotate4 (+-1) cuts; TE
             with even / odd y and x mirrors; point-like line source, flux
             line; all fields including the on-plane H rows
  media_3d   3D 22x20x18, npml 4, 120 steps, sigma_d, sigma_b and a Lorentz +
             Drude pole pair with masks, point source, flux plane, f32 + f64
  absorber_3d  3D 22x20x18, 120 steps, f32 + f64, boundary modes without a
             full CPML: (abs) matched-sigma absorber n_abs 4 with sigma_d,
             sigma_b and the pole pair; (mixed) npml 4 with the absorber on
             the z axis (n_abs 4, absorber_axes (2,)); (pec) neither (the
             absorber-path arithmetic in a PEC box); point source, flux plane
  tfsf_tm    ModeSourceTM (TFSF eigenmode source, h_masks) on a 2D TM guide:
             masks, the electric / magnetic amplitude sequences, fields and
             flux after 600 steps, f64
  sim3d      mp.Simulation in 3D (dimensions from a 3-vector cell): a point
             and a plane (finite aperture) GaussianSource, flux plane, f64
  raster     polygon coverage (triangle, rotated square, concave L, a polygon
             partly outside the grid, a clockwise pentagon, a 48-gon) on a
             40x36 grid, a three-polygon painter's scene (TM eps, Kottke TE
             pair, chi3), and the layout plugin's eps map of a two-polygon
             component
  shapes     mp.Simulation 2D TM cell with a Cylinder, Sphere, Prism, rotated
             Block and Ellipsoid: eps map, Ez and a flux after t = 12, f64
  sim_absorber  mp.Simulation with mp.Absorber layers: a 2D TM and a 2D TE
             cell (Absorber on x, PML on y) and a 3D cell (Absorber all
             round): Ez / Hz / Ey and a flux, f64
  sim_sym    mp.Simulation 2D cells with symmetries, f64: Mirror(Y) on a
             guide (decay stop at a point below the plane), Mirror(X, -1) +
             Mirror(Y) with a source pair, Rotate2, Rotate4 and a TE
             Mirror(Y, -1): full-cell field, fluxes (lines crossing and not
             crossing the planes) and probed points
  multisrc   mp.Simulation 2D TM cell with two sources (point and line, own
             pulses and amplitudes): Ez and a flux after t = 10, plain and
             with an at_every step function splitting the run, f64
"""
import hashlib
import json
import math
import os
import sys
import time

sys.dont_write_bytecode = True
REF = "/root/reference"
if REF not in sys.path:
    sys.path.insert(0, REF)

import numpy as np  # noqa: E402
import torch  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
torch.set_num_threads(8)


def _np(t):
    return t.detach().cpu().numpy().copy()


def save(name, **arrays):
    path = os.path.join(HERE, f"{name}.npz")
    out = {}
    for k, v in arrays.items():
        if isinstance(v, torch.Tensor):
            v = _np(v)
        out[k] = np.asarray(v)
    np.savez_compressed(path, **out)
    return path


def splitmix_uniform(n, seed):
    """Counter-based U[0,1) floats (splitmix64); the same generator lives in
    the package's synthetic-input module so the GPU box regenerates inputs."""
    out = np.empty(n, dtype=np.float64)
    mask = (1 << 64) - 1
    for i in range(n):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


# --------------------------------------------------------------------- pml
def case_pml():
    from meep_torch.core.pml import cpml_profiles
    combos = [(256, 16, 0.0), (255, 16, 0.5), (64, 8, 0.0), (63, 8, 0.5),
              (2048, 32, 0.0), (2047, 32, 0.5), (100, (10, 0), 0.0),
              (99, (0, 7), 0.5), (40, 6, 0.0), (39, 6, 0.5)]
    arrays = {}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        for ci, (n, npml, off) in enumerate(combos):
            dx = 0.05
            dt = 0.5 * dx
            b, a, k = cpml_profiles(n, npml, off, dt, dx, dtype=dtype)
            arrays[f"{tag}_{ci}_b"] = b
            arrays[f"{tag}_{ci}_a"] = a
            arrays[f"{tag}_{ci}_k"] = k
    meta = [[n, list(npml) if isinstance(npml, tuple) else npml, off]
            for n, npml, off in combos]
    arrays["combos_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    save("pml", **arrays)


# --------------------------------------------------------------- 2D helpers
def _eps2d(nx, ny, dtype):
    eps = torch.ones(nx, ny, dtype=torch.float64)
    eps[:, 20:28] = 12.0                       # horizontal guide
    x = torch.arange(nx, dtype=torch.float64)[:, None]
    y = torch.arange(ny, dtype=torch.float64)[None, :]
    blob = torch.exp(-((x - 40.0) ** 2 + (y - 12.0) ** 2) / (2 * 3.0 ** 2))
    eps = eps + 3.0 * blob
    return eps.to(dtype)


def case_tm_small():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import DFTFieldsMonitor, FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_tm import init_tm_state, step_tm

    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=64, ny=48, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        eps = _eps2d(64, 48, dtype)
        m = build_tm_coeffs(g, eps_inf=eps, npml=8)
        freqs = [0.3, 0.4, 0.5, 0.6, 0.7]
        out = {"eps": eps}
        for sname, src in (
                ("point", LineSource(g, ix=12, j0=24, j1=25,
                                     profile=GaussianPulse(0.5, 1.0))),
                ("line", LineSource(g, ix=12, j0=16, j1=32,
                                    profile=GaussianPulse(0.5, 1.0)))):
            out[f"{sname}_mask"] = src.mask
            out[f"{sname}_amp"] = np.array([src.amplitude(t) for t in range(400)])
            for ck in (0, 150):
                mons = [FluxMonitorTM(g, freqs, x_normal=True, idx=50, lo=10, hi=38),
                        FluxMonitorTM(g, freqs, x_normal=False, idx=36, lo=10, hi=54),
                        DFTFieldsMonitor(g, [0.45, 0.55], component="Ez"),
                        DFTFieldsMonitor(g, [0.5], component="Hx")]
                st, acc = run_tm(init_tm_state(g), m, 400, source=src,
                                 monitors=mons, checkpoint_every=ck)
                p = f"{sname}_ck{ck}"
                for f in st._fields:
                    v = getattr(st, f)
                    if isinstance(v, torch.Tensor):
                        out[f"{p}_{f}"] = v
                for i, a in enumerate(acc):
                    out[f"{p}_acc{i}"] = a
                out[f"{p}_fluxx"] = mons[0].flux(acc[0], g.dx)
                out[f"{p}_fluxy"] = mons[1].flux(acc[1], g.dx)
            # early states (debug aid): after 1 and 10 steps, point/line source
            st = init_tm_state(g)
            for t in range(10):
                st = step_tm(st, m, src.amplitude(t), src.mask)
                if t in (0, 9):
                    out[f"{sname}_s{t + 1}_Ez"] = st.Ez
                    out[f"{sname}_s{t + 1}_Hx"] = st.Hx
                    out[f"{sname}_s{t + 1}_Hy"] = st.Hy
        # a state that starts from a nonzero, inconsistent Ez/Dz pair
        st0 = init_tm_state(g)
        x = torch.arange(64, dtype=torch.float64)[:, None]
        y = torch.arange(48, dtype=torch.float64)[None, :]
        G = torch.exp(-((x - 30.0) ** 2 + (y - 24.0) ** 2) / (2 * 4.0 ** 2)).to(dtype)
        st0 = st0._replace(Ez=G.clone(), Dz=(1.5 * G).clone())
        st, _ = run_tm(st0, m, 50)
        out["blob_Ez0"] = G
        for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
            out[f"blob_{f}"] = getattr(st, f)
        save(f"tm_small_{tag}", **out)


def case_te_small():
    from meep_torch.core.coeffs import build_te_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY
    from meep_torch.core.run import run_te
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_te import init_te_state

    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=64, ny=48, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        eps = _eps2d(64, 48, dtype)
        m = build_te_coeffs(g, eps_inf=eps, npml=8)
        freqs = [0.3, 0.4, 0.5, 0.6, 0.7]
        src = LineSource(g, ix=12, j0=16, j1=32, profile=GaussianPulse(0.5, 1.0))
        src.mask = src.mask[:, :-1].contiguous()          # Jy on the Ey grid
        out = {"eps": eps, "mask": src.mask,
               "amp": np.array([src.amplitude(t) for t in range(400)])}
        for ck in (0, 150):
            mons = [FluxMonitorTE(g, freqs, ix=50, lo=10, hi=38),
                    FluxMonitorTEY(g, freqs, iy=36, lo=10, hi=54),
                    DFTFieldsMonitor(g, [0.45, 0.55], component="Hz"),
                    DFTFieldsMonitor(g, [0.5], component="Ey")]
            st, acc = run_te(init_te_state(g), m, 400, source=src, monitors=mons,
                             checkpoint_every=ck)
            p = f"ck{ck}"
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor):
                    out[f"{p}_{f}"] = v
            for i, a in enumerate(acc):
                out[f"{p}_acc{i}"] = a
            out[f"{p}_fluxx"] = mons[0].flux(acc[0], g.dx)
            out[f"{p}_fluxy"] = mons[1].flux(acc[1], g.dx)
        save(f"te_small_{tag}", **out)


def c1_inputs(dtype):
    """C1: 256x256, dx=0.05, Courant 0.5, eps=12 rows [123,133), npml 16,
    point Ez source at (24, 128), GaussianPulse(0.15, 0.1), flux line x=232,
    rows [96, 160), 50 freqs in [0.1, 0.2]."""
    eps = torch.ones(256, 256, dtype=dtype)
    eps[:, 123:133] = 12.0
    return eps


def case_c1():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_tm import init_tm_state

    timing = {}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=256, ny=256, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        eps = c1_inputs(dtype)
        m = build_tm_coeffs(g, eps_inf=eps, npml=16)
        src = LineSource(g, ix=24, j0=128, j1=129, profile=GaussianPulse(0.15, 0.1))
        freqs = np.linspace(0.1, 0.2, 50).tolist()
        mon = FluxMonitorTM(g, freqs, x_normal=True, idx=232, lo=96, hi=160)
        t0 = time.perf_counter()
        st, (acc,) = run_tm(init_tm_state(g), m, 2000, source=src, monitors=[mon])
        dtw = time.perf_counter() - t0
        timing[tag] = dtw
        save(f"c1_{tag}", eps=eps, Ez=st.Ez, Hx=st.Hx, Hy=st.Hy,
             acc=acc, flux=mon.flux(acc, g.dx), wall_s=np.array(dtw),
             threads=np.array(torch.get_num_threads()))
    return timing


# ----------------------------------------------------------------------- 3D
def _eps3d(nx, ny, nz, dtype):
    eps = torch.full((nx, ny, nz), 2.1, dtype=torch.float64)
    eps[:, :, nz // 2:] = 1.0
    eps[:, ny // 2 - 5:ny // 2 + 5, nz // 2 - 3:nz // 2 + 3] = 12.1
    return eps.to(dtype)


def case_3d_small():
    from meep_torch.core.monitors_3d import (DFTFieldsMonitor3D, FluxMonitor3D,
                                             PlaneSource, PointSource3D, run_3d)
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_3d import build_3d_coeffs, init_3d_state
    from meep_torch.core.stepper_3d_lean import build_lean, run_3d_lean

    nx, ny, nz, npml, dx, n = 32, 28, 24, 5, 0.05, 200
    prof = GaussianPulse(1.0, 2.0, cutoff=2.0)
    freqs = [0.8, 1.0, 1.2, 1.4]
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        eps = _eps3d(nx, ny, nz, dtype)
        c = build_3d_coeffs(nx, ny, nz, dx, eps=eps, npml=npml, dtype=dtype)
        out = {"eps": eps}
        srcs = [("point", PointSource3D(c.dt, "Ez", 8, ny // 2, nz // 2, prof))]
        if dtype == torch.float64:
            srcs.append(("plane", PlaneSource(nx, ny, nz, c.dt, "Ey", 0, 8, prof,
                                              lo=npml, hi=min(nx, ny, nz) - npml)))
        for sname, src in srcs:
            mons = [FluxMonitor3D(c.dt, freqs, normal="x", idx=24, ap=npml + 1,
                                  dtype=dtype),
                    FluxMonitor3D(c.dt, freqs, normal="z", idx=14, ap=npml + 1,
                                  dtype=dtype),
                    DFTFieldsMonitor3D(c.dt, [1.1], "Ez", plane=(2, 12),
                                       dtype=dtype),
                    DFTFieldsMonitor3D(c.dt, [1.0], "Hy", plane=(0, 20),
                                       dtype=dtype)]
            st, acc = run_3d(init_3d_state(nx, ny, nz, npml=npml, dtype=dtype),
                             c, n, source=src, monitors=mons)
            for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                out[f"{sname}_{f}"] = getattr(st, f)
            out[f"{sname}_psi_Ez_x"] = st.psi_Ez_x
            for i, a in enumerate(acc):
                out[f"{sname}_acc{i}"] = a
            out[f"{sname}_flux0"] = mons[0].flux(acc[0], dx)
            out[f"{sname}_flux1"] = mons[1].flux(acc[1], dx)
            out[f"{sname}_amp"] = np.array([src.amplitude(t) for t in range(n)])
        if dtype == torch.float32:
            cl = build_lean(nx, ny, nz, dx, eps, npml, dtype=dtype, device="cpu")
            src = PointSource3D(cl.dt, "Ez", 8, ny // 2, nz // 2, prof)
            mon = FluxMonitor3D(cl.dt, freqs, normal="x", idx=24, ap=npml + 1,
                                dtype=dtype)
            view, acc = run_3d_lean(cl, n, source=src, monitors=[mon])
            for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                out[f"lean_{f}"] = getattr(view, f)
            out["lean_acc0"] = acc[0]
        save(f"d3_small_{tag}", **out)


# ------------------------------------------------------------------ grads
def _blob_eps(theta, nx, ny, dtype, cx=24.0, cy=24.0, r=7.0, edge=1.0):
    x = torch.arange(nx, dtype=dtype)[:, None]
    y = torch.arange(ny, dtype=dtype)[None, :]
    dist = torch.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    disk = 0.5 * (1.0 - torch.tanh((dist - r) / edge))
    return 1.0 + theta * disk


def case_grad_tm():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_tm import init_tm_state

    for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        g = Grid(nx=48, ny=48, dx=0.05, dtype=dtype, device="cpu")
        eps = _blob_eps(0.8, 48, 48, dtype).detach().clone().requires_grad_(True)
        m = build_tm_coeffs(g, eps_inf=eps, npml=6)
        src = LineSource(g, ix=8, j0=18, j1=30, profile=GaussianPulse(1.0, 0.4),
                         amp_func=lambda y: 5.0)
        mon = FluxMonitorTM(g, [0.9, 1.0, 1.1], x_normal=True, idx=34, lo=16, hi=32)
        _, (acc,) = run_tm(init_tm_state(g), m, 1400, source=src, monitors=[mon],
                           checkpoint_every=200)
        J = mon.flux(acc, g.dx).sum()
        J.backward()
        save(f"grad_tm_{tag}", eps=eps.detach(), mask=src.mask, J=J.detach(),
             acc=acc.detach(), grad=eps.grad,
             amp=np.array([src.amplitude(t) for t in range(1400)]))


def case_grad_3d():
    from meep_torch.core.adjoint3d import run_3d_adjoint
    from meep_torch.core.monitors_3d import DFTFieldsMonitor3D, PointSource3D
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_3d import build_3d_coeffs, init_3d_state

    nx, ny, nz = 30, 24, 20
    for dtype, tag in ((torch.float64, "f64"),):
        eps = torch.full((nx, ny, nz), 1.5, dtype=dtype)
        eps[12:18, 9:15, 8:12] = 4.0
        eps = eps.requires_grad_(True)
        c = build_3d_coeffs(nx, ny, nz, 0.05, eps=eps, npml=6, dtype=dtype)
        src = PointSource3D(c.dt, "Ez", nx // 2, ny // 2, nz // 2 - 3,
                            GaussianPulse(0.6, 0.25))
        mon = DFTFieldsMonitor3D(c.dt, [0.6], "Ez", plane=(0, nx - 10), dtype=dtype)
        st0 = init_3d_state(nx, ny, nz, npml=6, dtype=dtype)

        def loss_fn(accs):
            E2 = accs[0][0, 0] ** 2 + accs[0][0, 1] ** 2
            return E2.sum()

        L, accs = run_3d_adjoint(c, 90, 16, src, [mon], loss_fn, st0)
        save(f"grad_3d_{tag}", eps=eps.detach(), L=np.array(L), acc=accs[0],
             grad=eps.grad)


# ---------------------------------------------------------------- filters
def case_filters():
    from meep_torch.adjoint.filters import conic_filter, gaussian_filter, tanh_projection
    from meep_torch.adjoint.problem import DesignRegion, MaterialGrid

    rho = torch.tensor(splitmix_uniform(32 * 24, 7).reshape(32, 24), dtype=torch.float64)
    cf = conic_filter(rho, 0.2, 0.05)
    gf = gaussian_filter(rho, 0.1, 0.05)
    tp = tanh_projection(cf, 8.0, 0.5)
    dr = DesignRegion(MaterialGrid((32, 24), 2.1, 12.1), ix0=10, ix1=42, iy0=5, iy1=29)
    base = torch.ones(60, 40, dtype=torch.float64)
    painted = dr.paint(base, tp)
    dr2 = DesignRegion(MaterialGrid((32, 24), 2.1, 12.1), ix0=4, ix1=52, iy0=3, iy1=39)
    painted2 = dr2.paint(base, tp)
    save("filters", rho=rho, conic=cf, gauss=gf, tanh=tp, painted=painted,
         painted2=painted2)


# ----------------------------------------------------------- golden plain
def case_golden_plain():
    import meep_torch.mp as mp
    out = {}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        sim = mp.Simulation(
            cell_size=mp.Vector3(10, 6), resolution=10,
            geometry=[mp.Block(center=mp.Vector3(), size=mp.Vector3(mp.inf, 1, mp.inf),
                               material=mp.Medium(epsilon=12))],
            sources=[mp.Source(mp.GaussianSource(0.15, fwidth=0.1), component=mp.Ez,
                               center=mp.Vector3(-3, 0), size=mp.Vector3(0, 2))],
            boundary_layers=[mp.PML(1.0)], dtype=dtype)
        fl = sim.add_flux(0.15, 0.1, 3, mp.FluxRegion(center=mp.Vector3(3, 0),
                                                      size=mp.Vector3(0, 4)))
        sim.run(until=40)
        out[f"{tag}_flux"] = np.array(mp.get_fluxes(fl))
        out[f"{tag}_freqs"] = np.array(mp.get_flux_freqs(fl))
        out[f"{tag}_eps"] = sim.get_array(component=mp.Dielectric)
        out[f"{tag}_Ez"] = sim.get_array(component=mp.Ez)
    save("golden_plain", **out)


# --------------------------------------------------------------- media pins
def _media_maps(nx, ny):
    x = torch.arange(nx, dtype=torch.float64)[:, None]
    y = torch.arange(ny, dtype=torch.float64)[None, :]
    eps = 1.0 + 2.0 * torch.exp(-((x - 30.0) ** 2 + (y - 18.0) ** 2) / 30.0)
    r = torch.tensor(splitmix_uniform(nx * ny, 7).reshape(nx, ny))
    inner = torch.zeros(nx, ny, dtype=torch.float64)
    inner[8:nx - 8, 8:ny - 8] = 1.0
    left = inner.clone()
    left[nx // 2:] = 0.0
    right = inner - left
    return {"eps": eps, "m1": (r > 0.5).double() * left, "m2": (r <= 0.5).double() * left,
            "chi2": 0.3 * right * (y > ny / 2).double(), "chi3": 0.05 * right * (y <= ny / 2).double(),
            "sigma_d": 2.0 * r * right, "sigma_b": 1.0 * (1.0 - r) * left, "mu": 1.0 + 0.5 * r * inner}


POLES2 = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]


def case_media_tm():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_tm import init_tm_state
    nx, ny, n = 48, 40, 240
    mp_ = _media_maps(nx, ny)
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=nx, ny=ny, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        cv = {k: v.to(dtype) for k, v in mp_.items()}
        out = {k: v for k, v in cv.items()}
        freqs = [0.6, 0.8, 1.0]
        for case, kw, drive in (
                ("all", dict(poles=POLES2, pole_masks=[cv["m1"], cv["m2"]], chi2=cv["chi2"], chi3=cv["chi3"],
                             sigma_d=cv["sigma_d"], sigma_b=cv["sigma_b"], mu=cv["mu"]), 20.0),
                ("nl", dict(chi2=cv["chi2"], chi3=cv["chi3"]), 60.0)):
            m = build_tm_coeffs(g, eps_inf=cv["eps"], npml=6, **kw)
            src = LineSource(g, ix=10, j0=8, j1=ny - 8, profile=GaussianPulse(0.8, 0.6, cutoff=3.0))
            src.mask = src.mask * drive
            mon = FluxMonitorTM(g, freqs, x_normal=True, idx=38, lo=7, hi=ny - 7)
            st, (acc,) = run_tm(init_tm_state(g, n_poles=len(kw.get("poles", ()))), m, n, source=src,
                                monitors=[mon])
            out[f"{case}_mask"] = src.mask
            out[f"{case}_amp"] = np.array([src.amplitude(t) for t in range(n)])
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"{case}_{f}"] = v
            out[f"{case}_acc"] = acc
        save(f"media_tm_{tag}", **out)


MPOLES2 = [(1.0, 0.2, 1.5), (0.7, 0.1, 0.8)]


class _HSource:
    """A run_tm source with an electric line mask and one magnetic (Hy)
    current mask, each with its own pulse (the h_masks protocol of
    core/run.py:17-23)."""

    def __init__(self, mask, amp, hy_mask, h_amp):
        self.mask, self._a, self.h_masks, self._h = mask, amp, {"src_hy_mask": hy_mask}, h_amp

    def amplitude(self, t):
        return self._a(t)

    def h_amplitude(self, t):
        return self._h(t)


def case_mag_tm():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_tm import init_tm_state
    nx, ny, n = 48, 40, 240
    mp_ = _media_maps(nx, ny)
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=nx, ny=ny, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        cv = {k: v.to(dtype) for k, v in mp_.items()}
        out = {k: v for k, v in cv.items()}
        mm = [cv["m2"] + 0.5 * cv["m1"], cv["m1"]]
        out["mm1"], out["mm2"] = mm
        line = LineSource(g, ix=10, j0=8, j1=ny - 8, profile=GaussianPulse(0.8, 0.6, cutoff=3.0))
        mon = FluxMonitorTM(g, [0.6, 0.8, 1.0], x_normal=True, idx=38, lo=7, hi=ny - 7)
        m = build_tm_coeffs(g, eps_inf=cv["eps"], npml=6, poles=POLES2, pole_masks=[cv["m1"], cv["m2"]],
                            chi3=cv["chi3"], m_poles=MPOLES2, m_pole_masks=mm)
        line.mask = line.mask * 20.0
        st, (acc,) = run_tm(init_tm_state(g, n_poles=2, n_mpoles=2), m, n, source=line, monitors=[mon])
        out["mag_mask"] = line.mask
        out["mag_amp"] = np.array([line.amplitude(t) for t in range(n)])
        for f in st._fields:
            v = getattr(st, f)
            if isinstance(v, torch.Tensor) and v.numel():
                out[f"mag_{f}"] = v
        out["mag_acc"] = acc
        hy = torch.zeros(nx - 1, ny, dtype=dtype)
        hy[14, 10:ny - 10] = 1.0 / 0.05
        hp = GaussianPulse(1.0, 0.8, cutoff=3.0)
        src = _HSource(line.mask, line.amplitude, hy, lambda t: 0.7 * hp.amplitude((t + 1.0) * g.dt))
        m = build_tm_coeffs(g, eps_inf=cv["eps"], npml=6, m_poles=MPOLES2, m_pole_masks=mm)
        st, (acc,) = run_tm(init_tm_state(g, n_mpoles=2), m, n, source=src, monitors=[mon])
        out["magh_hy_mask"] = hy
        out["magh_h_amp"] = np.array([src.h_amplitude(t) for t in range(n)])
        for f in st._fields:
            v = getattr(st, f)
            if isinstance(v, torch.Tensor) and v.numel():
                out[f"magh_{f}"] = v
        out["magh_acc"] = acc
        save(f"mag_tm_{tag}", **out)


def case_media_te():
    from meep_torch.core.coeffs import build_te_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTE
    from meep_torch.core.run import run_te
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_te import init_te_state
    nx, ny, n = 48, 40, 240
    mp_ = _media_maps(nx, ny)
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=nx, ny=ny, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        cv = {k: v.to(dtype) for k, v in mp_.items()}
        out = {k: v for k, v in cv.items()}
        freqs = [0.6, 0.8, 1.0]
        for case, kw, drive in (
                ("all", dict(poles=POLES2[:1], pole_masks=[cv["m1"]], poles_x=POLES2[1:], pole_masks_x=[cv["m2"]],
                             poles_y=[(1.6, 0.2, 0.7)], pole_masks_y=[cv["m2"]], chi2=cv["chi2"], chi3=cv["chi3"],
                             sigma_d=cv["sigma_d"], sigma_b=cv["sigma_b"], mu=cv["mu"]), 20.0),
                ("nl", dict(chi2=cv["chi2"], chi3=cv["chi3"]), 60.0)):
            m = build_te_coeffs(g, eps_inf=cv["eps"], npml=6, **kw)
            src = LineSource(g, ix=10, j0=8, j1=ny - 8, profile=GaussianPulse(0.8, 0.6, cutoff=3.0))
            src.mask = (src.mask[:, :-1] * drive).contiguous()
            mon = FluxMonitorTE(g, freqs, ix=38, lo=7, hi=ny - 7)
            st0 = init_te_state(g, n_poles=int(m.c1x.shape[0]), n_poles_y=int(m.c1y.shape[0]))
            st, (acc,) = run_te(st0, m, n, source=src, monitors=[mon])
            out[f"{case}_mask"] = src.mask
            out[f"{case}_amp"] = np.array([src.amplitude(t) for t in range(n)])
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"{case}_{f}"] = v
            out[f"{case}_acc"] = acc
        save(f"media_te_{tag}", **out)


def case_tensor_te():
    from meep_torch.core.coeffs import build_te_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTE
    from meep_torch.core.run import run_te
    from meep_torch.core.sources import GaussianPulse, LineSource
    from meep_torch.core.stepper_te import init_te_state
    nx, ny, n = 48, 40, 240
    mp_ = _media_maps(nx, ny)
    x = torch.arange(nx, dtype=torch.float64)[:, None]
    y = torch.arange(ny, dtype=torch.float64)[None, :]
    blob = torch.exp(-((x - 26.0) ** 2 + (y - 20.0) ** 2) / 60.0)
    ten = (mp_["eps"] + 1.5 * blob, 0.8 * blob, mp_["eps"] + 0.5 * blob)
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=nx, ny=ny, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        cv = {k: v.to(dtype) for k, v in mp_.items()}
        tv = tuple(t.to(dtype) for t in ten)
        out = {k: v for k, v in cv.items()}
        out["exx"], out["exy"], out["eyy"] = tv
        for case, kw, drive in (
                ("ten", dict(eps_tensor=tv, poles=POLES2[:1], pole_masks=[cv["m2"]],
                             poles_tensor=[(1.3, 0.2, (0.9, 0.4, 0.5))], pole_masks_tensor=[cv["m1"]]), 20.0),
                ("tenk", dict(eps_tensor=tv, chi3=cv["chi3"]), 60.0)):
            m = build_te_coeffs(g, eps_inf=cv["eps"], npml=6, **kw)
            src = LineSource(g, ix=10, j0=8, j1=ny - 8, profile=GaussianPulse(0.8, 0.6, cutoff=3.0))
            src.mask = (src.mask[:, :-1] * drive).contiguous()
            mon = FluxMonitorTE(g, [0.6, 0.8, 1.0], ix=38, lo=7, hi=ny - 7)
            st0 = init_te_state(g, n_poles=int(m.c1x.shape[0]), n_poles_y=int(m.c1y.shape[0]))
            st, (acc,) = run_te(st0, m, n, source=src, monitors=[mon])
            out[f"{case}_mask"] = src.mask
            out[f"{case}_amp"] = np.array([src.amplitude(t) for t in range(n)])
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"{case}_{f}"] = v
            out[f"{case}_acc"] = acc
        save(f"tensor_te_{tag}", **out)


FOLD_TM = {"ym+": dict(y_mirror=1), "ym-": dict(y_mirror=-1), "xm+": dict(x_mirror=1),
           "xm-": dict(x_mirror=-1), "xy": dict(x_mirror=-1, y_mirror=1), "r2+": dict(y_rot2=1),
           "r2-": dict(y_rot2=-1), "r4+": dict(rot4=1), "r4-": dict(rot4=-1)}
FOLD_TE = {"ym+": dict(y_mirror=1), "ym-": dict(y_mirror=-1), "xm+": dict(x_mirror=1),
           "xm-": dict(x_mirror=-1)}


def case_fold_2d():
    from meep_torch.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTE, FluxMonitorTM
    from meep_torch.core.run import run_te, run_tm
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_te import init_te_state
    from meep_torch.core.stepper_tm import init_tm_state
    n, steps = 36, 200
    x = torch.arange(n, dtype=torch.float64)[:, None]
    y = torch.arange(n, dtype=torch.float64)[None, :]
    eps = 1.0 + 1.5 * torch.exp(-((x - 14.0) ** 2 + (y - 11.0) ** 2) / 40.0)
    amp = GaussianPulse(0.9, 0.7, cutoff=3.0)

    class Src:
        def __init__(self, mask, dt):
            self.mask, self._dt = mask, dt

        def amplitude(self, t):
            return amp.amplitude((t + 0.5) * self._dt)

    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = Grid(nx=n, ny=n, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
        out = {"eps": eps.to(dtype)}
        for name, kw in FOLD_TM.items():
            m = build_tm_coeffs(g, eps_inf=eps.to(dtype), npml=((0 if kw.get("x_mirror") or kw.get("rot4") else 6, 6),
                                                                (0 if kw.get("y_mirror") or kw.get("y_rot2") or
                                                                 kw.get("rot4") else 6, 6)), **kw)
            mask = torch.zeros(n, n, dtype=dtype)
            mask[9, 4:14] = 20.0
            mask[4:12, 9] += 10.0
            st0 = init_tm_state(g, **kw)
            mon = FluxMonitorTM(g, [0.7, 0.9, 1.1], x_normal=True, idx=26, lo=0, hi=28)
            st, (acc,) = run_tm(st0, m, steps, source=Src(mask, g.dt), monitors=[mon])
            out[f"tm{name}_mask"] = mask
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"tm{name}_{f}"] = v
            out[f"tm{name}_acc"] = acc
        for name, kw in FOLD_TE.items():
            m = build_te_coeffs(g, eps_inf=eps.to(dtype), npml=((0 if kw.get("x_mirror") else 6, 6),
                                                                (0 if kw.get("y_mirror") else 6, 6)), **kw)
            mask = torch.zeros(n, n - 1, dtype=dtype)
            mask[9, 0:12] = 20.0
            mask[0:12, 9] += 10.0
            mon = FluxMonitorTE(g, [0.7, 0.9, 1.1], ix=26, lo=0, hi=28)
            st, (acc,) = run_te(init_te_state(g), m, steps, source=Src(mask, g.dt), monitors=[mon])
            out[f"te{name}_mask"] = mask
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"te{name}_{f}"] = v
            out[f"te{name}_acc"] = acc
        out["amp"] = np.array([Src(None, g.dt).amplitude(t) for t in range(steps)])
        save(f"fold_2d_{tag}", **out)


def case_media_3d():
    from meep_torch.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_3d import build_3d_coeffs, init_3d_state
    nx, ny, nz, npml, n = 22, 20, 18, 4, 120
    r = torch.tensor(splitmix_uniform(nx * ny * nz, 9).reshape(nx, ny, nz))
    inner = torch.zeros(nx, ny, nz, dtype=torch.float64)
    inner[npml + 2:nx - npml - 2, npml + 2:ny - npml - 2, npml + 2:nz - npml - 2] = 1.0
    eps = 1.0 + 2.0 * r * inner
    half = (r > 0.5).double() * inner
    maps = {"eps": eps, "sigma_d": 3.0 * r * inner, "sigma_b": 1.5 * (1.0 - r) * inner,
            "m1": half, "m2": inner - half}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        cv = {k: v.to(dtype) for k, v in maps.items()}
        c = build_3d_coeffs(nx, ny, nz, 0.05, eps=cv["eps"], npml=npml, sigma_d=cv["sigma_d"],
                            sigma_b=cv["sigma_b"], poles=POLES2, pole_masks=[cv["m1"], cv["m2"]],
                            dtype=dtype, device="cpu")
        src = PointSource3D(c.dt, "Ez", 7, ny // 2, nz // 2, GaussianPulse(1.0, 2.0, cutoff=2.0))
        mon = FluxMonitor3D(c.dt, [0.8, 1.0, 1.2], normal="x", idx=15, ap=npml + 1, dtype=dtype)
        st, (acc,) = run_3d(init_3d_state(nx, ny, nz, npml=npml, n_poles=2, dtype=dtype, device="cpu"), c, n,
                            source=src, monitors=[mon])
        out = dict(cv)
        out["amp"] = np.array([src.amplitude(t) for t in range(n)])
        for f in st._fields:
            v = getattr(st, f)
            if isinstance(v, torch.Tensor) and v.numel():
                out[f] = v
        out["acc"] = acc
        out["flux"] = mon.flux(acc, 0.05)
        save(f"media_3d_{tag}", **out)


def case_lean_poles_3d():
    """The lean in-place stepper with two ADE poles on masks
    (stepper_3d_lean.py:143-165,207-231): fields, the P(n+1) / P(n)
    buffers and a flux spectrum after 120 FP32 steps."""
    from meep_torch.core.monitors_3d import FluxMonitor3D, PointSource3D
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_3d_lean import build_lean, run_3d_lean
    nx, ny, nz, npml, n, dx = 22, 20, 18, 4, 120, 0.05
    r = torch.tensor(splitmix_uniform(nx * ny * nz, 33).reshape(nx, ny, nz))
    inner = torch.zeros(nx, ny, nz, dtype=torch.float64)
    inner[npml + 2:nx - npml - 2, npml + 2:ny - npml - 2, npml + 2:nz - npml - 2] = 1.0
    half = (r > 0.5).double() * inner
    dtype = torch.float32
    cv = {"eps": (1.0 + 2.0 * r * inner).to(dtype), "m1": half.to(dtype), "m2": (inner - half).to(dtype)}
    c = build_lean(nx, ny, nz, dx, cv["eps"], npml, poles=POLES2, pole_masks=[cv["m1"], cv["m2"]],
                   dtype=dtype, device="cpu")
    src = PointSource3D(c.dt, "Ez", 7, ny // 2, nz // 2, GaussianPulse(1.0, 2.0, cutoff=2.0))
    mon = FluxMonitor3D(c.dt, [0.8, 1.0, 1.2], normal="x", idx=15, ap=npml + 1, dtype=dtype)
    view, (acc,) = run_3d_lean(c, n, source=src, monitors=[mon])
    out = dict(cv)
    out["amp"] = np.array([src.amplitude(t) for t in range(n)])
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        out[f] = getattr(view, f)
    for key in ("x", "y", "z"):
        out[f"P{key}_new"], out[f"P{key}_old"] = c.P[key][0], c.P[key][1]
    out["acc"] = acc
    save("lean_poles_3d_f32", **out)


def case_absorber_3d():
    from meep_torch.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_3d import build_3d_coeffs, init_3d_state
    nx, ny, nz, n = 22, 20, 18, 120
    r = torch.tensor(splitmix_uniform(nx * ny * nz, 21).reshape(nx, ny, nz))
    inner = torch.zeros(nx, ny, nz, dtype=torch.float64)
    inner[6:nx - 6, 6:ny - 6, 6:nz - 6] = 1.0
    half = (r > 0.5).double() * inner
    maps = {"eps": 1.0 + 2.0 * r * inner, "sigma_d": 2.0 * r * inner, "sigma_b": 1.0 * (1.0 - r) * inner,
            "m1": half, "m2": inner - half}
    cases = {"abs": dict(npml=0, n_abs=4, media=True), "mixed": dict(npml=4, n_abs=4, axes=(2,), media=False),
             "pec": dict(npml=0, n_abs=0, media=False)}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        cv = {k: v.to(dtype) for k, v in maps.items()}
        out = dict(cv)
        for case, o in cases.items():
            kw = dict(sigma_d=cv["sigma_d"], sigma_b=cv["sigma_b"], poles=POLES2,
                      pole_masks=[cv["m1"], cv["m2"]]) if o["media"] else {}
            c = build_3d_coeffs(nx, ny, nz, 0.05, eps=cv["eps"], npml=o["npml"], n_abs=o["n_abs"],
                                absorber_axes=o.get("axes", ()), dtype=dtype, device="cpu", **kw)
            src = PointSource3D(c.dt, "Ez", 9, ny // 2, nz // 2, GaussianPulse(1.0, 2.0, cutoff=2.0))
            mon = FluxMonitor3D(c.dt, [0.8, 1.0, 1.2], normal="x", idx=14, ap=2, dtype=dtype)
            st0 = init_3d_state(nx, ny, nz, npml=o["npml"], n_poles=2 if o["media"] else 0, dtype=dtype,
                                device="cpu")
            st, (acc,) = run_3d(st0, c, n, source=src, monitors=[mon])
            for f in st._fields:
                v = getattr(st, f)
                if isinstance(v, torch.Tensor) and v.numel():
                    out[f"{case}_{f}"] = v
            out[f"{case}_acc"] = acc
            out[f"{case}_flux"] = mon.flux(acc, 0.05)
        out["amp"] = np.array([src.amplitude(t) for t in range(n)])
        save(f"absorber_3d_{tag}", **out)


def case_tfsf_tm():
    from meep_torch.core.coeffs import build_tm_coeffs
    from meep_torch.core.grid import Grid
    from meep_torch.core.monitors import FluxMonitorTM
    from meep_torch.core.run import run_tm
    from meep_torch.core.sources import GaussianPulse
    from meep_torch.core.stepper_tm import init_tm_state
    from meep_torch.ports.mode_source import ModeSourceTM
    from meep_torch.ports.port import Port
    nx, ny, n = 160, 64, 600
    dtype = torch.float64
    g = Grid(nx=nx, ny=ny, dx=0.05, courant=0.5, dtype=dtype, device="cpu")
    eps = torch.ones(nx, ny, dtype=dtype)
    eps[:, 28:36] = 12.0
    m = build_tm_coeffs(g, eps_inf=eps, npml=12)
    port = Port(axis="x", index=40, lo=14, hi=ny - 14, direction=+1)
    src = ModeSourceTM(g, m, port, GaussianPulse(0.6, 0.3), 0.6)
    mons = [FluxMonitorTM(g, [0.5, 0.6, 0.7], x_normal=True, idx=120, lo=14, hi=ny - 14),
            FluxMonitorTM(g, [0.5, 0.6, 0.7], x_normal=True, idx=20, lo=14, hi=ny - 14)]
    st, acc = run_tm(init_tm_state(g), m, n, source=src, monitors=mons)
    out = {"eps": eps, "mask": src.mask, "neff": np.array(src.neff),
           "amp": np.array([src.amplitude(t) for t in range(n)]),
           "hamp": np.array([src.h_amplitude(t) for t in range(n)])}
    for k, v in src.h_masks.items():
        out[k] = v
    for f in ("Ez", "Hx", "Hy", "Dz"):
        out[f] = getattr(st, f)
    for i, a in enumerate(acc):
        out[f"acc{i}"] = a
    save("tfsf_tm_f64", **out)


def case_sim3d():
    import meep_torch.mp as mp
    out = {}
    for kind in ("point", "plane"):
        cell = mp.Vector3(1.6, 1.4, 1.2)
        geom = [mp.Block(center=mp.Vector3(0.1, 0, -0.1), size=mp.Vector3(mp.inf, 0.3, 0.25),
                         material=mp.Medium(epsilon=6.0))]
        if kind == "point":
            srcs = [mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ez,
                              center=mp.Vector3(-0.4, 0.05, 0.1))]
        else:
            srcs = [mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ey,
                              center=mp.Vector3(-0.4, 0, 0), size=mp.Vector3(0, 0.6, 0.5))]
        sim = mp.Simulation(cell_size=cell, resolution=20, geometry=geom, sources=srcs,
                            boundary_layers=[mp.PML(0.2)], dimensions=3)
        fl = sim.add_flux(1.0, 0.4, 3, mp.FluxRegion(center=mp.Vector3(0.4, 0, 0), size=mp.Vector3(0, 1.4, 1.2)))
        sim.run(until=6.0)
        out[f"{kind}_flux"] = np.array(mp.get_fluxes(fl))
        for comp, name in ((mp.Ex, "Ex"), (mp.Ey, "Ey"), (mp.Ez, "Ez")):
            out[f"{kind}_{name}"] = np.asarray(sim.get_array(component=comp))
    save("sim3d_f64", **out)


# ------------------------------------------------------------- polygon raster
RASTER_POLYS = {
    "triangle": [[0.11, 0.23], [1.37, 0.41], [0.52, 1.63]],
    "rot_square": [[0.8, 0.05], [1.55, 0.8], [0.8, 1.55], [0.05, 0.8]],
    "concave_l": [[0.2, 0.2], [1.4, 0.2], [1.4, 0.6], [0.6, 0.6], [0.6, 1.5], [0.2, 1.5]],
    "outside": [[-0.3, 0.9], [1.1, -0.25], [2.4, 1.3], [0.7, 2.2]],
    "cw_pentagon": [[0.9, 1.6], [1.5, 1.1], [1.3, 0.3], [0.5, 0.3], [0.3, 1.1]],
}


def case_raster():
    from meep_torch.core.grid import Grid
    from meep_torch.geometry.polygons import Polygon
    from meep_torch.geometry.raster import coverage
    from meep_torch.geometry.scene import rasterize_scene
    g = Grid(nx=40, ny=36, dx=0.05, device="cpu")
    out = {}
    for name, v in RASTER_POLYS.items():
        out[f"cov_{name}"] = coverage(Polygon(np.array(v)), g)
    ring = np.stack([0.95 + 0.55 * np.cos(np.linspace(0, 2 * np.pi, 48, endpoint=False)),
                     0.9 + 0.55 * np.sin(np.linspace(0, 2 * np.pi, 48, endpoint=False))], axis=1)
    out["ngon"] = ring
    out["cov_ngon"] = coverage(Polygon(ring), g)
    items = [(Polygon(np.array(RASTER_POLYS["concave_l"])), 12.0), (Polygon(ring), 4.0),
             (Polygon(np.array(RASTER_POLYS["triangle"])), 2.25)]
    rm = rasterize_scene(items, g, background=1.44, chi3_items=[(Polygon(ring), 0.02)])
    out.update(scene_eps_tm=rm.eps_tm, scene_te_xx=rm.eps_te_xx, scene_te_yy=rm.eps_te_yy, scene_chi3=rm.chi3)
    # the layout plugin's steps (plugins/gdsfactory.py:19-42) on a duck-typed
    # component: a straight guide and a bent taper on layer 1, padded 0.4 um
    polys = [np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 0.5], [0.0, 0.5]]),
             np.array([[2.0, 0.0], [3.1, 0.4], [3.3, 1.2], [2.9, 1.25], [2.6, 0.7], [2.0, 0.5]])]
    res, pad, n_core, n_clad = 20, 0.4, 3.48, 1.44
    xs = np.concatenate([p[:, 0] for p in polys])
    ys = np.concatenate([p[:, 1] for p in polys])
    x0, y0 = xs.min() - pad, ys.min() - pad
    sx, sy = (xs.max() + pad) - x0, (ys.max() + pad) - y0
    pg = Grid(nx=int(round(sx * res)), ny=int(round(sy * res)), dx=1.0 / res, device="cpu")
    rm = rasterize_scene([(Polygon(p - [x0, y0]), n_core ** 2) for p in polys], pg, background=n_clad ** 2)
    out.update(plugin_eps=rm.eps_tm, plugin_nx=np.array(pg.nx), plugin_ny=np.array(pg.ny))
    save("raster", **out)


# ------------------------------------------------------- facade 2D shapes
def case_shapes():
    import meep_torch.mp as mp
    out = {}
    geom = [mp.Cylinder(radius=0.7, center=mp.Vector3(-1.5, 0.3), material=mp.Medium(epsilon=6.0)),
            mp.Sphere(radius=0.5, center=mp.Vector3(1.2, -0.8), material=mp.Medium(epsilon=3.0)),
            mp.Prism(vertices=[mp.Vector3(0, 0.5), mp.Vector3(1.5, 1.2), mp.Vector3(0.8, 1.8)],
                     material=mp.Medium(epsilon=9.0)),
            mp.Block(size=mp.Vector3(1.2, 0.4), center=mp.Vector3(-0.2, -1.0), e1=mp.Vector3(1, 1),
                     e2=mp.Vector3(-1, 1), material=mp.Medium(epsilon=4.0)),
            mp.Ellipsoid(size=mp.Vector3(1.6, 0.8), center=mp.Vector3(2.0, 1.0), material=mp.Medium(epsilon=2.5))]
    sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10, geometry=geom,
                        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=mp.Ez,
                                           center=mp.Vector3(-2.4, 0.1))],
                        boundary_layers=[mp.PML(0.5)])
    fl = sim.add_flux(0.8, 0.4, 3, mp.FluxRegion(center=mp.Vector3(2.6, 0), size=mp.Vector3(0, 3)))
    sim.run(until=12)
    out["eps"] = np.asarray(sim.get_array(component=mp.Dielectric))
    out["Ez"] = np.asarray(sim.get_array(component=mp.Ez))
    out["flux"] = np.array(mp.get_fluxes(fl))
    save("shapes_f64", **out)


# ---------------------------------------------------------- several sources
def case_multisrc():
    import meep_torch.mp as mp
    out = {}
    for hook in (False, True):
        sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10,
                            geometry=[mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))],
                            sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=mp.Ez,
                                               center=mp.Vector3(-2.0, 0.0)),
                                     mp.Source(mp.GaussianSource(1.1, fwidth=0.5), component=mp.Ez,
                                               center=mp.Vector3(-1.4, 0.5), size=mp.Vector3(0, 0.8),
                                               amplitude=-0.7)],
                            boundary_layers=[mp.PML(0.5)])
        fl = sim.add_flux(0.9, 0.5, 4, mp.FluxRegion(center=mp.Vector3(2.2, 0), size=mp.Vector3(0, 3)))
        seen = []
        if hook:
            sim.run(mp.at_every(1.5, lambda s: seen.append(float(s.get_field_point(mp.Ez, mp.Vector3(1.0, 0))))),
                    until=10)
        else:
            sim.run(until=10)
        tag = "hook" if hook else "plain"
        out[f"{tag}_Ez"] = np.asarray(sim.get_array(component=mp.Ez))
        out[f"{tag}_flux"] = np.array(mp.get_fluxes(fl))
        if hook:
            out["hook_probe"] = np.array(seen)
    save("multisrc_f64", **out)


def case_sim_absorber():
    import meep_torch.mp as mp
    out = {}
    for pol, comp in (("tm", mp.Ez), ("te", mp.Hz)):
        sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10,
                            geometry=[mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))],
                            sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=comp,
                                               center=mp.Vector3(-1.8, 0.0), size=mp.Vector3(0, 0.8))],
                            boundary_layers=[mp.Absorber(0.8, direction=mp.X), mp.PML(0.5, direction=mp.Y)])
        fl = sim.add_flux(0.8, 0.5, 4, mp.FluxRegion(center=mp.Vector3(1.6, 0), size=mp.Vector3(0, 3)))
        sim.run(until=12)
        out[f"{pol}_field"] = np.asarray(sim.get_array(component=comp))
        out[f"{pol}_flux"] = np.array(mp.get_fluxes(fl))
    sim = mp.Simulation(cell_size=mp.Vector3(1.6, 1.4, 1.2), resolution=20,
                        geometry=[mp.Block(center=mp.Vector3(0.1, 0, -0.1), size=mp.Vector3(mp.inf, 0.3, 0.25),
                                           material=mp.Medium(epsilon=6.0))],
                        sources=[mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ey,
                                           center=mp.Vector3(-0.4, 0, 0), size=mp.Vector3(0, 0.6, 0.5))],
                        boundary_layers=[mp.Absorber(0.25)], dimensions=3)
    fl = sim.add_flux(1.0, 0.4, 3, mp.FluxRegion(center=mp.Vector3(0.4, 0, 0), size=mp.Vector3(0, 1.0, 0.8)))
    sim.run(until=6.0)
    out["d3_flux"] = np.array(mp.get_fluxes(fl))
    out["d3_Ey"] = np.asarray(sim.get_array(component=mp.Ey))
    save("sim_absorber_f64", **out)


def _sym_cases(mp):
    """(name, Simulation kwargs, flux regions, probe points, run kwargs) of the
    symmetry golden; also used by the tests through the recorded values."""
    guide = [mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))]
    pulse = lambda f: mp.GaussianSource(f, fwidth=0.6)
    return [
        ("my", dict(cell_size=mp.Vector3(6, 4), geometry=guide,
                    sources=[mp.Source(pulse(0.8), component=mp.Ez, center=mp.Vector3(-2.0, 0.0),
                                       size=mp.Vector3(0, 0.8))],
                    symmetries=[mp.Mirror(mp.Y)]),
         [mp.FluxRegion(center=mp.Vector3(2.0, 0), size=mp.Vector3(0, 3)),
          mp.FluxRegion(center=mp.Vector3(0.5, -1.0), size=mp.Vector3(2.0, 0))],
         [(1.0, 0.3), (1.0, -0.3)], dict(until_after_sources=mp.stop_when_fields_decayed(2.0, mp.Ez,
                                                                                        mp.Vector3(1.0, -0.3), 1e-2))),
        ("mxy", dict(cell_size=mp.Vector3(5, 4), geometry=[],
                     sources=[mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(0.8, 0.0)),
                              mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(-0.8, 0.0), amplitude=-1.0)],
                     symmetries=[mp.Mirror(mp.X, phase=-1), mp.Mirror(mp.Y)]),
         [mp.FluxRegion(center=mp.Vector3(0, 1.2), size=mp.Vector3(4.0, 0)),
          mp.FluxRegion(center=mp.Vector3(1.8, 0), size=mp.Vector3(0, 3.0))],
         [(0.5, 0.5), (-0.5, -0.5), (-1.2, 0.7)], dict(until=8)),
        ("r2", dict(cell_size=mp.Vector3(4, 4), geometry=[mp.Block(size=mp.Vector3(1.2, 0.4),
                                                                   material=mp.Medium(epsilon=6.0))],
                    sources=[mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(0.5, 0.5)),
                             mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(-0.5, -0.5))],
                    symmetries=[mp.Rotate2(mp.Z)]),
         [mp.FluxRegion(center=mp.Vector3(1.2, 0.9), size=mp.Vector3(0, 1.4))],
         [(0.6, 0.4), (-0.6, -0.4), (0.9, -0.8)], dict(until=8)),
        ("r4", dict(cell_size=mp.Vector3(4, 4), geometry=[mp.Block(size=mp.Vector3(0.8, 0.8),
                                                                   material=mp.Medium(epsilon=4.0))],
                    sources=[mp.Source(pulse(1.0), component=mp.Ez, center=mp.Vector3(0.6, 0.7))],
                    symmetries=[mp.Rotate4(mp.Z)]),
         [mp.FluxRegion(center=mp.Vector3(1.4, 0.9), size=mp.Vector3(0, 1.4))],
         [(0.6, 0.7), (-0.7, 0.6), (-0.6, -0.7)], dict(until=8)),
        ("te", dict(cell_size=mp.Vector3(6, 4), geometry=guide,
                    sources=[mp.Source(pulse(0.8), component=mp.Ey, center=mp.Vector3(-2.0, 0.0),
                                       size=mp.Vector3(0, 0.8))],
                    symmetries=[mp.Mirror(mp.Y, phase=-1)]),
         [mp.FluxRegion(center=mp.Vector3(2.0, 0), size=mp.Vector3(0, 3))],
         [(1.0, 0.3)], dict(until=10)),
    ]


def case_sim_sym():
    import meep_torch.mp as mp
    out = {}
    for name, kw, regions, pts, runkw in _sym_cases(mp):
        sim = mp.Simulation(resolution=10, boundary_layers=[mp.PML(0.5)], **kw)
        fls = [sim.add_flux(0.8, 0.4, 3, r) for r in regions]
        seen = []
        comp = mp.Ez if name != "te" else mp.Hz
        sim.run(mp.at_every(1.0, lambda s: seen.append([float(np.real(s.get_field_point(comp, mp.Vector3(*p))))
                                                          for p in pts])), **runkw)
        out[f"{name}_field"] = np.asarray(sim.get_array(component=comp))
        out[f"{name}_flux"] = np.array([mp.get_fluxes(f) for f in fls])
        out[f"{name}_probe"] = np.array(seen)
        out[f"{name}_t"] = np.array(sim.meep_time())
    save("sim_sym_f64", **out)


def main():
    which = sys.argv[1:] or ["pml", "tm", "te", "c1", "3d", "grad_tm", "grad_3d",
                             "filters", "golden_plain", "media_tm", "media_te", "media_3d",
                             "absorber_3d", "tfsf_tm", "sim3d", "raster", "shapes", "multisrc", "sim_absorber",
                             "mag_tm", "tensor_te", "fold_2d", "sim_sym", "lean_poles_3d"]
    t_all = time.perf_counter()
    timing = {}
    try:
        with open(os.path.join(HERE, "MANIFEST.json")) as fh:
            timing = json.load(fh).get("timing", {})
    except (OSError, ValueError):
        pass
    for w in which:
        t0 = time.perf_counter()
        {"pml": case_pml, "tm": case_tm_small, "te": case_te_small,
         "c1": lambda: timing.update(c1=case_c1()), "3d": case_3d_small,
         "grad_tm": case_grad_tm, "grad_3d": case_grad_3d,
         "filters": case_filters, "golden_plain": case_golden_plain,
         "media_tm": case_media_tm, "media_te": case_media_te, "media_3d": case_media_3d,
         "absorber_3d": case_absorber_3d,
         "lean_poles_3d": case_lean_poles_3d,
         "tfsf_tm": case_tfsf_tm, "sim3d": case_sim3d, "raster": case_raster,
         "shapes": case_shapes, "multisrc": case_multisrc, "sim_absorber": case_sim_absorber,
         "mag_tm": case_mag_tm, "tensor_te": case_tensor_te, "fold_2d": case_fold_2d, "sim_sym": case_sim_sym}[w]()
        print(f"[golden] {w}: {time.perf_counter() - t0:.1f}s", flush=True)
    man = {}
    for fn in sorted(os.listdir(HERE)):
        if fn.endswith(".npz"):
            with open(os.path.join(HERE, fn), "rb") as fh:
                man[fn] = hashlib.sha256(fh.read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump({"sha256": man, "torch": torch.__version__,
                   "threads": torch.get_num_threads(), "timing": timing}, fh, indent=1)
    print(f"[golden] total {time.perf_counter() - t_all:.1f}s")


if __name__ == "__main__":
    main()
