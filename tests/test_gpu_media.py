"""TM media on the HIP engine (the two-pass kernels): electric ADE poles
(Lorentz and Drude), the chi2 / chi3 fixed point and sigma_d / sigma_b / mu
(meep_torch/core/coeffs.py:87-106,154-200, stepper_tm.py:116-121,161-163,
198-231).

* every term at once against the oracle: FP64 fields to 1e-12, FP32 within
  the 1e-5 fields / 1e-4 spectra bars, with and without checkpoint chunks,
  and one step of step_tm from a random state;
* the reference's physics scenes through run_tm: a Lorentz slab against the
  transfer matrix (3%, meep_torch tests/analytic/test_dispersion.py) and a
  sigma_d slab against meep's transmission (amplitude rms 3%, the fixture of
  meep_torch tests/examples/test_w5_conductivity.py);
* eps gradients through conductive media (sigma_d, sigma_b, mu) and ADE
  poles, TM and TE, against CPU autograd through the oracle; gradients
  through Kerr media, and the batched launch, refuse loudly."""
import cmath
import json
import os

import numpy as np
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]
HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, NPML, DX = 96, 64, 10, 0.05
FREQS = [0.6, 0.8, 1.0]


def _media(dtype):
    from meep_torch_amd import synth
    r = torch.tensor(synth.splitmix_uniform(NX * NY, 31).reshape(NX, NY), dtype=torch.float64)
    blk = torch.zeros(NX, NY, dtype=torch.float64)
    blk[50:75, 16:48] = 1.0
    eps = 1.0 + 1.5 * blk * r
    kw = dict(poles=[(1.2, 0.1, 1.5), (0.0, 0.3, 0.5, 0.9)],        # Lorentz, Drude
              pole_masks=[blk.to(dtype), (blk * (r > 0.5)).to(dtype)],
              chi3=(0.05 * blk).to(dtype), chi2=(0.02 * blk * r).to(dtype),
              sigma_d=(0.3 * blk).to(dtype), sigma_b=(0.1 * blk).to(dtype), mu=(1.0 + 0.5 * blk).to(dtype))
    return eps.to(dtype), kw


def _hip(dtype, ck=0, n=400):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    eps, kw = _media(dtype)
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=eps.cuda(), npml=NPML,
                        **{k: ([x.cuda() for x in v] if k == "pole_masks" else v.cuda() if torch.is_tensor(v) else v)
                           for k, v in kw.items()})
    # drive so that chi3 |E|^2 reaches about 0.07 in the block: far above
    # rounding, mildly nonlinear
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5), amp_func=lambda y: 5.0)
    mon = FluxMonitorTM(g, FREQS, x_normal=True, idx=80, lo=12, hi=52)
    st, (acc,) = run_tm(init_tm_state(g, n_poles=2), m, n, source=src, monitors=[mon], checkpoint_every=ck)
    return st, acc, src


def _oracle(dtype, src, n=400):
    from oracle import yee2d
    torch.set_num_threads(8)
    eps, kw = _media(dtype)
    c = yee2d.tm_coeffs(NX, NY, DX, eps, NPML, dtype=dtype, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 80, "lo": 12, "hi": 52, "freqs": FREQS}]
    amps = [src.amplitude(t) for t in range(n)]
    return yee2d.run2d("tm", yee2d.tm_state(NX, NY, dtype, n_poles=2), c, n, amps=amps,
                       mask=src.mask.cpu(), monitors=mons)


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("dtype,tf,ts", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)],
                         ids=["f64", "f32"])
@pytest.mark.parametrize("ck", [0, 150])
def test_media_run_matches_oracle(dtype, tf, ts, ck):
    st, acc, src = _hip(dtype, ck)
    s, (oacc,) = _oracle(dtype, src)
    names = ("Ez", "Dz", "Hx", "Hy", "P_cur", "P_prev")
    errs = {f: rel(getattr(st, f), s[f]) for f in names}
    errs["acc"] = rel(acc, oacc)
    print(f"\n[media {dtype} ck={ck}] " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert float(s["P_cur"].abs().max()) > 0 and float(s["Ez"].abs().max()) > 0
    for f in names:
        assert errs[f] <= tf, (f, errs[f])
    assert errs["acc"] <= ts


def test_media_step_tm_matches_oracle():
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.stepper_tm import init_tm_state, step_tm
    from oracle import yee2d
    dt = torch.float64
    eps, kw = _media(dt)
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    m = build_tm_coeffs(g, eps_inf=eps.cuda(), npml=NPML,
                        **{k: ([x.cuda() for x in v] if k == "pole_masks" else v.cuda() if torch.is_tensor(v) else v)
                           for k, v in kw.items()})
    gen = torch.Generator().manual_seed(4)
    st = init_tm_state(g, n_poles=2)
    st = st._replace(**{f: torch.randn(getattr(st, f).shape, generator=gen, dtype=dt).cuda()
                        for f in ("Ez", "Dz", "Hx", "Hy", "P_cur", "P_prev")})
    mask = torch.zeros(NX, NY, dtype=dt)
    mask[20, 16:48] = 1.0 / DX
    out = step_tm(st, m, 0.7, mask.cuda())
    c = yee2d.tm_coeffs(NX, NY, DX, eps, NPML, dtype=dt, **kw)
    ref = yee2d.tm_step({f: getattr(st, f).cpu() for f in ("Ez", "Dz", "Hx", "Hy", "P_cur", "P_prev", "psi_Ezy",
                                                          "psi_Ezx", "psi_Hyx", "psi_Hxy")},
                        c, 0.7, mask)                 # raw amplitude: sigma_d scales it by d2
    for f in ("Ez", "Dz", "Hx", "Hy", "P_cur", "P_prev"):
        assert rel(getattr(out, f), ref[f]) <= 1e-13, f


def test_lorentz_slab_transmission_on_hip():
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    pole, eps_inf = (2.0, 0.1, 1.5), 2.25
    freqs = [0.50, 0.55, 0.60, 0.65, 0.70, 0.75, 0.80]

    def flux(on):
        g = Grid(nx=1400, ny=8, dx=0.02, dtype=torch.float64, device="cuda")
        slab = torch.zeros(1400, 8, dtype=torch.float64, device="cuda")
        slab[600:700, :] = 1.0
        kw = dict(eps_inf=1.0 + (eps_inf - 1.0) * slab, poles=[pole], pole_masks=[slab]) if on else {}
        m = build_tm_coeffs(g, npml=40, y_pmc=True, **kw)
        src = LineSource(g, ix=100, j0=0, j1=8, profile=GaussianPulse(0.65, 0.35))
        mon = FluxMonitorTM(g, freqs, x_normal=True, idx=1200, lo=0, hi=8)
        _, (acc,) = run_tm(init_tm_state(g, n_poles=1 if on else 0, y_pmc=True), m, 9000, source=src,
                           monitors=[mon])
        return mon.flux(acc, g.dx).cpu()

    T = flux(True) / flux(False)
    worst = 0.0
    for i, f in enumerate(freqs):
        fn, gm, s = pole
        n2 = cmath.sqrt(eps_inf + s * fn * fn / (fn * fn - f * f - 1j * f * gm))
        delta = 2 * cmath.pi * f * n2 * 2.0
        r12 = (1 - n2) / (1 + n2)
        t = (2 / (1 + n2)) * (2 * n2 / (1 + n2)) * cmath.exp(1j * delta) / (1 - r12 ** 2 * cmath.exp(2j * delta))
        e = abs(float(T[i]) - abs(t) ** 2) / abs(t) ** 2
        worst = max(worst, e)
        assert e < 0.03, (f, float(T[i]), abs(t) ** 2)
    print(f"\n[lorentz slab] worst relative error vs transfer matrix {worst:.2e}")


def test_conductive_slab_vs_meep_on_hip():
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    with open(os.path.join(HERE, "golden", "meep_conductive_slab.json")) as fh:
        ref = json.load(fh)
    dx = 1.0 / ref["resolution"]
    nx = int(round(30 / dx))

    def flux(on):
        g = Grid(nx=nx, ny=8, dx=dx, dtype=torch.float32, device="cuda")
        sd = torch.zeros(nx, 8, dtype=torch.float32, device="cuda")
        sd[int(round(14 / dx)):int(round(16 / dx)), :] = ref["sigma_d"]
        m = build_tm_coeffs(g, npml=int(round(1.5 / dx)), y_pmc=True, **({"sigma_d": sd} if on else {}))
        src = LineSource(g, ix=int(round(5 / dx)), j0=0, j1=8, profile=GaussianPulse(ref["fcen"], ref["df"]))
        mon = FluxMonitorTM(g, ref["freqs"], x_normal=True, idx=int(round(21 / dx)), lo=0, hi=8)
        _, (acc,) = run_tm(init_tm_state(g, y_pmc=True), m, 10000, source=src, monitors=[mon])
        return mon.flux(acc, g.dx).double().cpu().numpy()

    amp = np.sqrt(flux(True) / flux(False))
    rms = float(np.sqrt(np.mean((amp / np.sqrt(ref["T_power"]) - 1.0) ** 2)))
    print(f"\n[conductive slab fp32] amplitude rms vs meep {rms:.2e}")
    assert rms < 0.03


def test_media_refuse_gradients_and_batches():
    """TE gradients with respect to chi3 itself (only eps_x / eps_y are
    differentiated) raise; so does graph capture of a batch whose variants
    carry media (they run one by one)."""
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    g = Grid(nx=32, ny=32, dx=0.05, dtype=torch.float64, device="cuda")
    chi3 = torch.full((32, 32), 0.01, dtype=torch.float64, device="cuda", requires_grad=True)
    m = build_te_coeffs(g, eps_inf=2.0, npml=4, chi3=chi3)
    with pytest.raises(NotImplementedError, match="coefficient"):
        run_te(init_te_state(g), m, 5)
    m2 = build_tm_coeffs(g, eps_inf=2.0, npml=4, poles=[(1.0, 0.1, 1.0)])
    with pytest.raises(NotImplementedError, match="graph"):
        BatchedTM([(m2, None, [])], 10).capture()


def _te_media(dtype):
    from meep_torch_amd import synth
    r = torch.tensor(synth.splitmix_uniform(NX * NY, 37).reshape(NX, NY), dtype=torch.float64)
    blk = torch.zeros(NX, NY, dtype=torch.float64)
    blk[50:75, 16:48] = 1.0
    eps = 1.0 + 1.5 * blk * r
    kw = dict(poles=[(1.2, 0.1, 1.5)], pole_masks=[blk.to(dtype)],
              poles_x=[(0.0, 0.3, 0.5, 0.9)], pole_masks_x=[(blk * (r > 0.5)).to(dtype)],      # Drude on Ex only
              sigma_d=(0.3 * blk).to(dtype), sigma_b=(0.1 * blk).to(dtype), mu=(1.0 + 0.5 * blk).to(dtype),
              chi3=(0.05 * blk).to(dtype), chi2=(0.02 * blk * r).to(dtype))
    return eps.to(dtype), kw


@pytest.mark.parametrize("dtype,tf,ts", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)],
                         ids=["f64", "f32"])
def test_te_media_run_matches_oracle(dtype, tf, ts):
    """TE: a pole shared by Ex and Ey, a Drude pole on Ex only, sigma_d,
    sigma_b, mu, chi2 and chi3 (three Jacobi passes coupling Ex and Ey),
    against the oracle's TE step."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    eps, kw = _te_media(dtype)
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dtype, device="cuda")
    dev = {k: ([x.cuda() for x in v] if k.startswith("pole_masks") else v.cuda() if torch.is_tensor(v) else v)
           for k, v in kw.items()}
    m = build_te_coeffs(g, eps_inf=eps.cuda(), npml=NPML, **dev)
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5), amp_func=lambda y: 5.0)
    src.mask = src.mask[:, :-1].contiguous()
    mon = FluxMonitorTE(g, FREQS, ix=80, lo=12, hi=52)
    st, (acc,) = run_te(init_te_state(g, n_poles=2, n_poles_y=1), m, 400, source=src, monitors=[mon],
                        checkpoint_every=150)
    torch.set_num_threads(8)
    c = yee2d.te_coeffs(NX, NY, DX, eps, NPML, dtype=dtype, **kw)
    s, (oacc,) = yee2d.run2d("te", yee2d.te_state(NX, NY, dtype, npx=2, npy=1), c, 400,
                             amps=[src.amplitude(t) for t in range(400)], mask=src.mask.cpu(),
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 80, "lo": 12, "hi": 52,
                                        "freqs": FREQS}], checkpoint_every=150)
    names = ("Hz", "Ex", "Ey", "Dx", "Dy", "Px_cur", "Px_prev", "Py_cur", "Py_prev")
    errs = {f: rel(getattr(st, f), s[f]) for f in names}
    errs["acc"] = rel(acc, oacc)
    print(f"\n[te media {dtype}] " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert float(s["Px_cur"][1].abs().max()) > 0
    for f in names:
        assert errs[f] <= tf, (f, errs[f])
    assert errs["acc"] <= ts


def test_te_lorentz_slab_on_hip():
    """Parallel-plate TE (PEC y walls, PML along x): the Lorentz slab within
    3% of the transfer matrix through run_te."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    pole, eps_inf = (2.0, 0.1, 1.5), 2.25
    freqs = [0.50, 0.60, 0.70, 0.80]

    def flux(on):
        g = Grid(nx=1400, ny=8, dx=0.02, dtype=torch.float64, device="cuda")
        slab = torch.zeros(1400, 8, dtype=torch.float64, device="cuda")
        slab[600:700, :] = 1.0
        kw = dict(eps_inf=1.0 + (eps_inf - 1.0) * slab, poles=[pole], pole_masks=[slab]) if on else {}
        m = build_te_coeffs(g, npml=((40, 40), (0, 0)), **kw)
        src = LineSource(g, ix=100, j0=0, j1=8, profile=GaussianPulse(0.65, 0.35))
        src.mask = src.mask[:, :-1].contiguous()
        mon = FluxMonitorTE(g, freqs, ix=1200, lo=0, hi=7)
        _, (acc,) = run_te(init_te_state(g, n_poles=1 if on else 0), m, 9000, source=src, monitors=[mon])
        return mon.flux(acc, g.dx).cpu()

    T = flux(True) / flux(False)
    for i, f in enumerate(freqs):
        fn, gm, s = pole
        n2 = cmath.sqrt(eps_inf + s * fn * fn / (fn * fn - f * f - 1j * f * gm))
        delta = 2 * cmath.pi * f * n2 * 2.0
        r12 = (1 - n2) / (1 + n2)
        t = (2 / (1 + n2)) * (2 * n2 / (1 + n2)) * cmath.exp(1j * delta) / (1 - r12 ** 2 * cmath.exp(2j * delta))
        assert abs(float(T[i]) - abs(t) ** 2) / abs(t) ** 2 < 0.03, (f, float(T[i]), abs(t) ** 2)


def test_te_facade_drude_block_equals_direct_run():
    """A TE Simulation cell with a cell-aligned Drude + D-conductivity + chi3
    block against run_te on hand-built maps (FP64): identical fields and flux."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    D = mp.DrudeSusceptibility(frequency=0.9, gamma=0.3, sigma=0.5)
    sim = mp.Simulation(
        cell_size=mp.Vector3(10, 6), resolution=10,
        geometry=[mp.Block(center=mp.Vector3(1.5, 0), size=mp.Vector3(1, 2),
                           material=mp.Medium(epsilon=2, E_susceptibilities=[D], D_conductivity=0.2,
                                              chi3=0.05))],
        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.5), component=mp.Ey,
                           center=mp.Vector3(-4, 0), size=mp.Vector3(0, 2))],
        boundary_layers=[mp.PML(1.0)], dtype=torch.float64, device="cuda")
    fl = sim.add_flux(0.8, 0.4, 3, mp.FluxRegion(center=mp.Vector3(4, 0), size=mp.Vector3(0, 4)))
    sim.run(until=20)
    g = sim.grid
    z = lambda v=0.0: torch.full((100, 60), v, dtype=torch.float64, device="cuda")
    eps, mD, sd, x3 = z(1.0), z(), z(), z()
    eps[60:70, 20:40] = 2.0
    mD[60:70, 20:40] = 1.0
    sd[60:70, 20:40] = 0.2
    x3[60:70, 20:40] = 0.05
    m = build_te_coeffs(g, eps_inf=eps, npml=((10, 10), (10, 10)), poles=[D.to_pole()], pole_masks=[mD],
                        sigma_d=sd, sigma_b=z(), mu=z(1.0), chi3=x3)
    src = sim._make_sources()[0]
    freqs = np.linspace(0.8 - 0.4 / 2.0, 0.8 + 0.4 / 2.0, 3).tolist()
    kind, mon, meta = sim._monitors[fl.name]
    mon2 = FluxMonitorTE(g, freqs, ix=meta["idx"], lo=meta["lo"], hi=meta["hi"])
    st, (acc,) = run_te(init_te_state(g, n_poles=1), m, sim._t_step, source=src, monitors=[mon2])
    assert np.array_equal(sim.get_array(component=mp.Ey), st.Ey.cpu().numpy())
    assert np.array_equal(np.asarray(mp.get_fluxes(fl)), mon2.flux(acc, g.dx).cpu().numpy())
    assert float(st.Px_cur.abs().max()) > 0


@pytest.mark.parametrize("ck", [0, 100])
def test_tm_eps_gradient_through_conductive_media(ck):
    """The hand-written adjoint transposes the conductivity recurrences
    (Dz' = d1 Dz + d2 curl, H' = h1 H -+ h2 curl): the eps gradient of a flux
    objective through a sigma_d / sigma_b / mu block equals CPU autograd
    through the oracle's steps (FP64), with and without checkpoint chunks."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _media(dt)
    kw = {k: kw0[k] for k in ("sigma_d", "sigma_b", "mu")}
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    eps = eps0.clone().cuda().requires_grad_(True)
    m = build_tm_coeffs(g, eps_inf=eps, npml=NPML, **{k: v.cuda() for k, v in kw.items()})
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5))
    mon = FluxMonitorTM(g, FREQS, x_normal=True, idx=80, lo=12, hi=52)
    _, (acc,) = run_tm(init_tm_state(g), m, n, source=src, monitors=[mon], checkpoint_every=ck)
    mon.flux(acc, DX).sum().backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.tm_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 80, "lo": 12, "hi": 52, "freqs": FREQS}]
    _, oacc = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dt), c, n, amps=[src.amplitude(t) for t in range(n)],
                          mask=src.mask.cpu(), monitors=mons)
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc[0], DX).sum(), e)
    err = rel(eps.grad, g_ref)
    print(f"\n[cond grad ck={ck}] rel-L2 {err:.2e}")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0 and err <= 1e-9


def test_te_eps_gradient_through_conductive_media():
    """TE counterpart: the adjoint's Dx / Dy / Hz conductivity transposes
    give the eps gradient of a flux objective through a sigma_d / sigma_b /
    mu block equal to CPU autograd through the oracle's TE steps (FP64)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _te_media(dt)
    kw = {k: kw0[k] for k in ("sigma_d", "sigma_b", "mu")}
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    eps = eps0.clone().cuda().requires_grad_(True)
    m = build_te_coeffs(g, eps_inf=eps, npml=NPML, **{k: v.cuda() for k, v in kw.items()})
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5))
    src.mask = src.mask[:, :-1].contiguous()
    mon = FluxMonitorTE(g, FREQS, ix=80, lo=12, hi=52)
    _, (acc,) = run_te(init_te_state(g), m, n, source=src, monitors=[mon], checkpoint_every=100)
    mon.flux(acc, DX).sum().backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.te_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    _, (oacc,) = yee2d.run2d("te", yee2d.te_state(NX, NY, dt), c, n, amps=[src.amplitude(t) for t in range(n)],
                             mask=src.mask.cpu(), monitors=[{"type": "flux", "kind": "te_x", "idx": 80, "lo": 12,
                                                             "hi": 52, "freqs": FREQS}])
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("te_x", oacc, DX).sum(), e)
    err = rel(eps.grad, g_ref)
    print(f"\n[te cond grad] rel-L2 {err:.2e}")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0 and err <= 1e-9


@pytest.mark.parametrize("cond", [False, True], ids=["poles", "poles+sigma"])
def test_tm_eps_gradient_through_poles(cond):
    """The TM adjoint transposes the ADE pole recurrence (P' = c1 P + c2 P_prev
    + c3 Ez, Ez' = (Dz' - sum P') / eps) with the pole stacks' cotangents
    carried between steps: the eps gradient of a flux objective through a
    Lorentz + Drude block (alone and with sigma_d / sigma_b / mu) equals CPU
    autograd through the oracle's steps (FP64), across checkpoint chunks."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _media(dt)
    names = ("poles", "pole_masks") + (("sigma_d", "sigma_b", "mu") if cond else ())
    kw = {k: kw0[k] for k in names}
    dev = {k: ([x.cuda() for x in v] if k == "pole_masks" else v.cuda() if torch.is_tensor(v) else v)
           for k, v in kw.items()}
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    eps = eps0.clone().cuda().requires_grad_(True)
    m = build_tm_coeffs(g, eps_inf=eps, npml=NPML, **dev)
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5))
    mon = FluxMonitorTM(g, FREQS, x_normal=True, idx=80, lo=12, hi=52)
    _, (acc,) = run_tm(init_tm_state(g, n_poles=2), m, n, source=src, monitors=[mon], checkpoint_every=100)
    mon.flux(acc, DX).sum().backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.tm_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 80, "lo": 12, "hi": 52, "freqs": FREQS}]
    _, oacc = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dt, n_poles=2), c, n,
                          amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(), monitors=mons)
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc[0], DX).sum(), e)
    err = rel(eps.grad, g_ref)
    print(f"\n[pole grad cond={cond}] rel-L2 {err:.2e}")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0 and err <= 1e-9


def test_te_eps_gradient_through_poles():
    """TE poles (one shared by Ex and Ey, a Drude pole on Ex only) with
    conductivity: the TE adjoint carries the Px / Py stacks' cotangents, and
    the eps gradient of a flux objective equals CPU autograd through the
    oracle's TE steps (FP64)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _te_media(dt)
    kw = {k: v for k, v in kw0.items() if k not in ("chi3", "chi2")}
    dev = {k: ([x.cuda() for x in v] if k.startswith("pole_masks") else v.cuda() if torch.is_tensor(v) else v)
           for k, v in kw.items()}
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    eps = eps0.clone().cuda().requires_grad_(True)
    m = build_te_coeffs(g, eps_inf=eps, npml=NPML, **dev)
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5))
    src.mask = src.mask[:, :-1].contiguous()
    mon = FluxMonitorTE(g, FREQS, ix=80, lo=12, hi=52)
    _, (acc,) = run_te(init_te_state(g, n_poles=2, n_poles_y=1), m, n, source=src, monitors=[mon],
                       checkpoint_every=100)
    mon.flux(acc, DX).sum().backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.te_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    _, (oacc,) = yee2d.run2d("te", yee2d.te_state(NX, NY, dt, npx=2, npy=1), c, n,
                             amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(),
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 80, "lo": 12, "hi": 52,
                                        "freqs": FREQS}])
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("te_x", oacc, DX).sum(), e)
    err = rel(eps.grad, g_ref)
    print(f"\n[te pole grad] rel-L2 {err:.2e}")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0 and err <= 1e-9


@pytest.mark.parametrize("with_poles", [False, True], ids=["kerr", "kerr+poles+sigma"])
def test_te_eps_gradient_through_kerr(with_poles):
    """The TE adjoint transposes the chi2/chi3 fixed point: the tape keeps
    rhs = D - sum P, the three Jacobi passes (coupled through the four-point
    averages of the other component) are recomputed from it and transposed
    last to first, the averages' cotangents scattered back with the other
    average.  The eps gradient of a flux objective equals CPU autograd
    through the oracle's TE steps (FP64), with chi2 and chi3 alone and with
    poles and conductivity in the same cells."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _te_media(dt)
    keep = ("chi3", "chi2") + (("poles", "pole_masks", "poles_x", "pole_masks_x", "sigma_d", "sigma_b", "mu")
                               if with_poles else ())
    kw = {k: v for k, v in kw0.items() if k in keep}
    kw["chi3"] = kw["chi3"] * 40.0      # a strong nonlinearity at this drive
    dev = {k: ([x.cuda() for x in v] if k.startswith("pole_masks") else v.cuda() if torch.is_tensor(v) else v)
           for k, v in kw.items()}
    npx, npy = (2, 1) if with_poles else (0, 0)
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    eps = eps0.clone().cuda().requires_grad_(True)
    m = build_te_coeffs(g, eps_inf=eps, npml=NPML, **dev)
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5), amp_func=lambda y: 30.0)
    src.mask = src.mask[:, :-1].contiguous()
    mon = FluxMonitorTE(g, FREQS, ix=80, lo=12, hi=52)
    _, (acc,) = run_te(init_te_state(g, n_poles=npx, n_poles_y=npy), m, n, source=src, monitors=[mon],
                       checkpoint_every=100)
    mon.flux(acc, DX).sum().backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.te_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    _, (oacc,) = yee2d.run2d("te", yee2d.te_state(NX, NY, dt, npx=npx, npy=npy), c, n,
                             amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(),
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 80, "lo": 12, "hi": 52,
                                        "freqs": FREQS}])
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("te_x", oacc, DX).sum(), e)
    err = rel(eps.grad, g_ref)
    # the fixed point matters here: dropping it changes the gradient by far more than the bar
    c_lin = yee2d.te_coeffs(NX, NY, DX, e, NPML, dtype=dt, **{k: v for k, v in kw.items() if k not in ("chi3", "chi2")})
    _, (lacc,) = yee2d.run2d("te", yee2d.te_state(NX, NY, dt, npx=npx, npy=npy), c_lin, n,
                             amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(),
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 80, "lo": 12, "hi": 52,
                                        "freqs": FREQS}])
    g_lin, = torch.autograd.grad(yee2d.flux_from_acc("te_x", lacc, DX).sum(), e)
    print(f"\n[te kerr grad] rel-L2 {err:.2e} (linear-media gradient differs by {rel(g_lin, g_ref):.2e})")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0 and err <= 1e-9
    assert float(rel(g_lin, g_ref)) > 1e-4


@pytest.mark.parametrize("with_poles", [False, True], ids=["kerr", "kerr+poles+sigma"])
def test_tm_eps_gradient_through_kerr(with_poles):
    """The TM adjoint transposes the chi2/chi3 fixed point (three passes of
    Ez = rhs / (eps + chi2 Ez + chi3 Ez^2) from rhs / eps; the tape keeps
    rhs = Dz - sum P and the iterates are recomputed from it): the eps
    gradient of a flux objective through a chi2 + chi3 block, driven hard
    enough that chi3 |E|^2 reaches about 0.07, equals CPU autograd through
    the oracle's steps (FP64; alone, and with the two poles and
    sigma_d / sigma_b / mu), across checkpoint chunks, and central finite
    differences of the HIP forward."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _media(dt)
    names = ("chi2", "chi3") + (("poles", "pole_masks", "sigma_d", "sigma_b", "mu") if with_poles else ())
    kw = {k: kw0[k] for k in names}
    npo = 2 if with_poles else 0
    dev = {k: ([x.cuda() for x in v] if k == "pole_masks" else v.cuda() if torch.is_tensor(v) else v)
           for k, v in kw.items()}
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5), amp_func=lambda y: 5.0)
    mon = FluxMonitorTM(g, FREQS, x_normal=True, idx=80, lo=12, hi=52)

    def hip_flux(e, ck=100):
        m = build_tm_coeffs(g, eps_inf=e, npml=NPML, **dev)
        _, (acc,) = run_tm(init_tm_state(g, n_poles=npo), m, n, source=src, monitors=[mon], checkpoint_every=ck)
        return mon.flux(acc, DX).sum()

    eps = eps0.clone().cuda().requires_grad_(True)
    hip_flux(eps).backward()
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    c = yee2d.tm_coeffs(NX, NY, DX, e, NPML, dtype=dt, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 80, "lo": 12, "hi": 52, "freqs": FREQS}]
    _, oacc = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dt, n_poles=npo), c, n,
                          amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(), monitors=mons)
    g_ref, = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc[0], DX).sum(), e)
    err = rel(eps.grad, g_ref)
    # the nonlinear terms matter: the gradient without them differs by far more
    m_lin = {k: v for k, v in kw.items() if k not in ("chi2", "chi3")}
    e2 = eps0.clone().requires_grad_(True)
    c2 = yee2d.tm_coeffs(NX, NY, DX, e2, NPML, dtype=dt, **m_lin)
    _, oacc2 = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dt, n_poles=npo), c2, n,
                           amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(), monitors=mons)
    g_lin, = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc2[0], DX).sum(), e2)
    gap = rel(g_lin, g_ref)
    fd = []
    h = 1e-5
    for (i, j) in ((60, 30), (70, 40), (55, 20)):
        ep, em = eps0.clone().cuda(), eps0.clone().cuda()
        ep[i, j] += h
        em[i, j] -= h
        with torch.no_grad():
            d = (float(hip_flux(ep, 0)) - float(hip_flux(em, 0))) / (2 * h)
        fd.append(abs(float(eps.grad[i, j]) - d) / max(abs(d), 1e-30))
    print(f"\n[kerr grad poles={with_poles}] rel-L2 vs oracle autograd {err:.2e} (linear-model gradient "
          f"differs by {gap:.2e}); FD rel errors {[f'{x:.1e}' for x in fd]}")
    assert float(g_ref[50:75, 16:48].abs().max()) > 0
    assert gap > 1e6 * err      # measured: 5.7e-4 (2.0e-5 with the damping media) against 5e-16 to 9e-16
    assert err <= 1e-9
    assert max(fd) <= 1e-5


def _gc_loss(eps=None, chi3=None, pole_mask=None, n_steps=10, checkpoint_every=4, chi3_fixed=None):
    """The reference's gradcheck scene (tests/autodiff/test_gradcheck.py:15-27):
    16 x 16, 3-cell PML, a Gaussian blob in Ez and Dz, loss = sum Ez^2."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=16, ny=16, dx=0.05, dtype=torch.float64, device="cuda")
    x = torch.arange(16, dtype=torch.float64, device="cuda")[:, None]
    y = torch.arange(16, dtype=torch.float64, device="cuda")[None, :]
    blob = torch.exp(-((x - 8.0) ** 2 + (y - 8.0) ** 2) / 4.0)
    poles = [(1.5, 0.2, 0.8)] if pole_mask is not None else []
    masks = [pole_mask] if pole_mask is not None else None
    m = build_tm_coeffs(g, eps_inf=eps, poles=poles, pole_masks=masks,
                        chi3=chi3 if chi3 is not None else chi3_fixed, npml=3)
    s = init_tm_state(g, n_poles=len(poles))._replace(Ez=blob.clone(), Dz=blob.clone())
    s, _ = run_tm(s, m, n_steps, checkpoint_every=checkpoint_every)
    return (s.Ez ** 2).sum()


def test_gradcheck_chi3_and_pole_mask():
    """Gradients with respect to the chi3 map and a Lorentz pole's mask (the
    reference's own gradcheck cases) pass torch.autograd.gradcheck through
    the HIP adjoint (FP64, checkpoints every 4 steps), and the pole-mask
    gradient under chi3 (where the adjoint recomputes each step's starting
    Ez from the taped Dz - sum P) equals central differences."""
    x0 = torch.full((16, 16), 0.05, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda c: _gc_loss(eps=2.0, chi3=c), (x0,), eps=1e-6, atol=1e-4)
    m0 = torch.full((16, 16), 0.5, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda mk: _gc_loss(eps=2.0, pole_mask=mk), (m0,), eps=1e-6, atol=1e-4)
    chi = torch.full((16, 16), 0.05, dtype=torch.float64, device="cuda")
    assert torch.autograd.gradcheck(lambda mk: _gc_loss(eps=2.0, pole_mask=mk, chi3_fixed=chi), (m0,), eps=1e-6,
                                    atol=1e-4)
    e0 = torch.full((16, 16), 2.0, dtype=torch.float64, device="cuda", requires_grad=True)
    l1 = _gc_loss(eps=e0, checkpoint_every=0)
    g1, = torch.autograd.grad(l1, e0)
    l2 = _gc_loss(eps=e0, checkpoint_every=3)
    g2, = torch.autograd.grad(l2, e0)
    assert torch.allclose(g1, g2, atol=1e-12, rtol=1e-10)


def test_chi3_and_mask_gradients_match_oracle_autograd():
    """On the 96 x 64 media scene (chi2 + chi3 block, driven so chi3 |E|^2
    reaches about 0.07, with the Lorentz + Drude masks): the gradients of a
    flux objective with respect to chi3, chi2 and both pole masks equal CPU
    autograd through the oracle's steps (FP64), across checkpoint chunks."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    from oracle import yee2d
    dt = torch.float64
    eps0, kw0 = _media(dt)
    n = 300
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    src = LineSource(g, ix=20, j0=16, j1=48, profile=GaussianPulse(0.8, 0.5), amp_func=lambda y: 5.0)
    mon = FluxMonitorTM(g, FREQS, x_normal=True, idx=80, lo=12, hi=52)
    leaves = {"chi3": kw0["chi3"], "chi2": kw0["chi2"], "m0": kw0["pole_masks"][0], "m1": kw0["pole_masks"][1]}
    hl = {k: v.clone().cuda().requires_grad_(True) for k, v in leaves.items()}
    m = build_tm_coeffs(g, eps_inf=eps0.cuda(), npml=NPML, poles=kw0["poles"], pole_masks=[hl["m0"], hl["m1"]],
                        chi3=hl["chi3"], chi2=hl["chi2"])
    _, (acc,) = run_tm(init_tm_state(g, n_poles=2), m, n, source=src, monitors=[mon], checkpoint_every=100)
    gh = torch.autograd.grad(mon.flux(acc, DX).sum(), list(hl.values()))
    torch.set_num_threads(8)
    cl = {k: v.clone().requires_grad_(True) for k, v in leaves.items()}
    c = yee2d.tm_coeffs(NX, NY, DX, eps0, NPML, dtype=dt, poles=kw0["poles"], pole_masks=[cl["m0"], cl["m1"]],
                        chi3=cl["chi3"], chi2=cl["chi2"])
    mons = [{"type": "flux", "kind": "tm_x", "idx": 80, "lo": 12, "hi": 52, "freqs": FREQS}]
    _, oacc = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dt, n_poles=2), c, n,
                          amps=[src.amplitude(t) for t in range(n)], mask=src.mask.cpu(), monitors=mons)
    go = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc[0], DX).sum(), list(cl.values()))
    errs = {k: rel(a, b) for k, a, b in zip(hl, gh, go)}
    print("\n[coefficient grads] " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for k, b in zip(hl, go):
        assert float(b.abs().max()) > 0, k
    for k, v in errs.items():
        assert v <= 1e-9, (k, v)
