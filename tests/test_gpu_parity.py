"""Parity of the HIP engine against the reference (golden vectors written by the
reference) and the CPU oracle.  Bars (tolerances written here):
  fields   relative L2 <= 1e-5 vs the reference's FP32 stepper (FP64: 1e-12)
  spectra  relative L2 <= 1e-4 (FP64: 1e-10)
The printed lines also state the error of the FP32 engine against the FP64
reference run."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

FREQS5 = [0.3, 0.4, 0.5, 0.6, 0.7]
TOL = {torch.float32: (1e-5, 1e-4), torch.float64: (1e-12, 1e-10)}


def rel(a, b):
    a = a.detach().double().cpu()
    b = torch.as_tensor(np.asarray(b)).double()
    assert tuple(a.shape) == tuple(b.shape), (tuple(a.shape), tuple(b.shape))
    n = float(b.norm())
    return float((a - b).norm()) / (n if n > 0 else 1.0)


def _src_mask(g, sname):
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    if sname == "point":
        return LineSource(g, ix=12, j0=24, j1=25, profile=GaussianPulse(0.5, 1.0))
    return LineSource(g, ix=12, j0=16, j1=32, profile=GaussianPulse(0.5, 1.0))


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("sname", ["point", "line"])
@pytest.mark.parametrize("ck", [0, 150])
def test_tm_small(tag, dtype, sname, ck):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import DFTFieldsMonitor, FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    G = golden(f"tm_small_{tag}")
    g = Grid(nx=64, ny=48, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=torch.from_numpy(G["eps"]).cuda(), npml=8)
    src = _src_mask(g, sname)
    assert torch.equal(src.mask.cpu(), torch.from_numpy(G[f"{sname}_mask"]))
    mons = [FluxMonitorTM(g, FREQS5, x_normal=True, idx=50, lo=10, hi=38),
            FluxMonitorTM(g, FREQS5, x_normal=False, idx=36, lo=10, hi=54),
            DFTFieldsMonitor(g, [0.45, 0.55], component="Ez"),
            DFTFieldsMonitor(g, [0.5], component="Hx")]
    st, acc = run_tm(init_tm_state(g), m, 400, source=src, monitors=mons, checkpoint_every=ck)
    p = f"{sname}_ck{ck}"
    tf, ts = TOL[dtype]
    for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
        e = rel(getattr(st, f), G[f"{p}_{f}"])
        assert e <= tf, (f, e)
    for i in range(4):
        e = rel(acc[i], G[f"{p}_acc{i}"])
        assert e <= ts, (i, e)
    assert rel(mons[0].flux(acc[0], g.dx), G[f"{p}_fluxx"]) <= ts
    assert rel(mons[1].flux(acc[1], g.dx), G[f"{p}_fluxy"]) <= ts
    if dtype == torch.float32:
        G64 = golden("tm_small_f64")
        print(f"TM {p}: Ez vs ref-f32 {rel(st.Ez, G[f'{p}_Ez']):.2e}, "
              f"vs ref-f64 {rel(st.Ez, G64[f'{p}_Ez']):.2e}; flux vs f64 "
              f"{rel(mons[0].flux(acc[0], g.dx), G64[f'{p}_fluxx']):.2e}")


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_inconsistent_state(tag, dtype):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    G = golden(f"tm_small_{tag}")
    g = Grid(nx=64, ny=48, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=torch.from_numpy(G["eps"]).cuda(), npml=8)
    e0 = torch.from_numpy(G["blob_Ez0"]).cuda()
    s0 = init_tm_state(g)._replace(Ez=e0.clone(), Dz=1.5 * e0)
    st, _ = run_tm(s0, m, 50)
    for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
        assert rel(getattr(st, f), G[f"blob_{f}"]) <= TOL[dtype][0], f
    assert torch.equal(s0.Ez, e0), "run_tm must not modify its input state"


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("ck", [0, 150])
def test_te_small(tag, dtype, ck):
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    G = golden(f"te_small_{tag}")
    g = Grid(nx=64, ny=48, dx=0.05, dtype=dtype, device="cuda")
    m = build_te_coeffs(g, eps_inf=torch.from_numpy(G["eps"]).cuda(), npml=8)
    src = LineSource(g, ix=12, j0=16, j1=32, profile=GaussianPulse(0.5, 1.0))
    src.mask = src.mask[:, :-1].contiguous()
    mons = [FluxMonitorTE(g, FREQS5, ix=50, lo=10, hi=38),
            FluxMonitorTEY(g, FREQS5, iy=36, lo=10, hi=54),
            DFTFieldsMonitor(g, [0.45, 0.55], component="Hz"),
            DFTFieldsMonitor(g, [0.5], component="Ey")]
    st, acc = run_te(init_te_state(g), m, 400, source=src, monitors=mons, checkpoint_every=ck)
    p = f"ck{ck}"
    tf, ts = TOL[dtype]
    for f in ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"):
        assert rel(getattr(st, f), G[f"{p}_{f}"]) <= tf, f
    for i in range(4):
        assert rel(acc[i], G[f"{p}_acc{i}"]) <= ts, i
    assert rel(mons[0].flux(acc[0], g.dx), G[f"{p}_fluxx"]) <= ts
    assert rel(mons[1].flux(acc[1], g.dx), G[f"{p}_fluxy"]) <= ts


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_c1_full(tag, dtype):
    """C1 at full size (256x256, 16-cell PML, 2000 steps, 50-frequency flux)."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    G = golden(f"c1_{tag}")
    g = Grid(nx=256, ny=256, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=torch.from_numpy(G["eps"]).cuda(), npml=16)
    src = LineSource(g, ix=24, j0=128, j1=129, profile=GaussianPulse(0.15, 0.1))
    mon = FluxMonitorTM(g, np.linspace(0.1, 0.2, 50).tolist(), x_normal=True, idx=232, lo=96, hi=160)
    st, (acc,) = run_tm(init_tm_state(g), m, 2000, source=src, monitors=[mon])
    tf, ts = TOL[dtype]
    for f in ("Ez", "Hx", "Hy"):
        assert rel(getattr(st, f), G[f]) <= tf, f
    assert rel(acc, G["acc"]) <= ts
    fl = mon.flux(acc, g.dx)
    assert rel(fl, G["flux"]) <= ts
    if dtype == torch.float32:
        G64 = golden("c1_f64")
        print(f"C1 f32: Ez vs ref-f32 {rel(st.Ez, G['Ez']):.2e} vs ref-f64 {rel(st.Ez, G64['Ez']):.2e}; "
              f"flux vs ref-f32 {rel(fl, G['flux']):.2e} vs ref-f64 {rel(fl, G64['flux']):.2e}")


def _d3(tag, dtype, sname, lean=False):
    from meep_torch_amd.core.monitors_3d import (DFTFieldsMonitor3D, FluxMonitor3D, PlaneSource,
                                                 PointSource3D, run_3d)
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    from meep_torch_amd.core.stepper_3d_lean import build_lean, run_3d_lean
    G = golden(f"d3_small_{tag}")
    nx, ny, nz, npml, dx, n = 32, 28, 24, 5, 0.05, 200
    eps = torch.from_numpy(G["eps"]).cuda()
    prof = GaussianPulse(1.0, 2.0, cutoff=2.0)
    freqs = [0.8, 1.0, 1.2, 1.4]
    if lean:
        cl = build_lean(nx, ny, nz, dx, eps, npml, dtype=dtype, device="cuda")
        src = PointSource3D(cl.dt, "Ez", 8, ny // 2, nz // 2, prof)
        mon = FluxMonitor3D(cl.dt, freqs, normal="x", idx=24, ap=npml + 1, dtype=dtype)
        view, acc = run_3d_lean(cl, n, source=src, monitors=[mon])
        return G, view, acc, None
    c = build_3d_coeffs(nx, ny, nz, dx, eps=eps, npml=npml, dtype=dtype, device="cuda")
    if sname == "point":
        src = PointSource3D(c.dt, "Ez", 8, ny // 2, nz // 2, prof)
    else:
        src = PlaneSource(nx, ny, nz, c.dt, "Ey", 0, 8, prof, lo=npml,
                          hi=min(nx, ny, nz) - npml, device="cuda")
    mons = [FluxMonitor3D(c.dt, freqs, normal="x", idx=24, ap=npml + 1, dtype=dtype),
            FluxMonitor3D(c.dt, freqs, normal="z", idx=14, ap=npml + 1, dtype=dtype),
            DFTFieldsMonitor3D(c.dt, [1.1], "Ez", plane=(2, 12), dtype=dtype),
            DFTFieldsMonitor3D(c.dt, [1.0], "Hy", plane=(0, 20), dtype=dtype)]
    st, acc = run_3d(init_3d_state(nx, ny, nz, npml=npml, dtype=dtype, device="cuda"), c, n,
                     source=src, monitors=mons)
    return G, st, acc, mons


@pytest.mark.parametrize("tag,dtype,sname", [("f32", torch.float32, "point"),
                                             ("f64", torch.float64, "point"),
                                             ("f64", torch.float64, "plane")])
def test_3d_small(tag, dtype, sname):
    G, st, acc, mons = _d3(tag, dtype, sname)
    tf, ts = TOL[dtype]
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "psi_Ez_x"):
        assert rel(getattr(st, f), G[f"{sname}_{f}"]) <= tf, f
    for i in range(4):
        assert rel(acc[i], G[f"{sname}_acc{i}"]) <= ts, i
    assert rel(mons[0].flux(acc[0], 0.05), G[f"{sname}_flux0"]) <= ts
    assert rel(mons[1].flux(acc[1], 0.05), G[f"{sname}_flux1"]) <= ts


def test_3d_lean():
    G, view, acc, _ = _d3("f32", torch.float32, "point", lean=True)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert rel(getattr(view, f), G[f"lean_{f}"]) <= 1e-5, f
    assert rel(acc[0], G["lean_acc0"]) <= 1e-4


def test_bitwise_determinism():
    """Two identical runs give identical bits (no atomics / races)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    g = Grid(nx=300, ny=260, dx=0.05, dtype=torch.float32, device="cuda")
    m = build_te_coeffs(g, eps_inf=2.0, npml=20)
    src = LineSource(g, ix=40, j0=30, j1=230, profile=GaussianPulse(0.5, 0.3))
    src.mask = src.mask[:, :-1].contiguous()
    mon = FluxMonitorTE(g, [0.4, 0.5, 0.6], ix=200, lo=30, hi=230)
    a = run_te(init_te_state(g), m, 300, source=src, monitors=[mon])
    b = run_te(init_te_state(g), m, 300, source=src, monitors=[mon])
    assert torch.equal(a[0].Hz, b[0].Hz) and torch.equal(a[1][0], b[1][0])
