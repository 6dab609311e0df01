"""TFSF magnetic currents (the h_masks / h_amplitude source protocol,
meep_torch/core/run.py:17-23, stepper_tm.py:169-193, stepper_te.py:120-128)
on the HIP engine against the oracle.

The source is a Huygens pair in a uniform medium with the reference's TFSF
timing (ports/mode_source.py): an electric line current plus a magnetic line
current half a cell upstream.  Checked: fields and flux spectra against the
oracle's run (FP64 to 1e-12, FP32 within the 1e-5 / 1e-4 bars), chunked runs,
the one-way launch on the HIP path, the single-step wrappers step_tm / step_te
with magnetic masks, and eps gradients with a TFSF source against CPU
autograd through the oracle."""
import math

import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, NPML, DX, N = 120, 64, 12, 0.05, 700
FREQS = [0.8, 1.0, 1.2]


def pulse(t, fc=1.0, fw=0.6, cutoff=5.0):
    w = 1.0 / fw
    t0 = cutoff * w
    return math.cos(2 * math.pi * fc * (t - t0)) * math.exp(-((t - t0) ** 2) / (2 * w * w))


class Huygens:
    """A line current on Dz / Dy at x index ix and a magnetic line current on
    Hy / Hz at ix - 1, with the reference's TFSF amplitudes."""

    def __init__(self, mode, dt, dtype, device):
        lo, hi = (NPML, NY - NPML) if mode == "tm" else (NPML, NY - 1 - NPML)
        shape = (NX, NY) if mode == "tm" else (NX, NY - 1)
        hshape = (NX - 1, NY) if mode == "tm" else (NX - 1, NY - 1)
        m = torch.zeros(shape, dtype=torch.float64)
        m[60, lo:hi] = 1.0 / DX
        h = torch.zeros(hshape, dtype=torch.float64)
        h[59, lo:hi] = 1.0 / DX
        self.mask = m.to(dtype).to(device)
        self.h_masks = {("src_hy_mask" if mode == "tm" else "src_hz_mask"): h.to(dtype).to(device)}
        self.mode, self.dt = mode, dt
        self.lo, self.hi = lo, hi

    def amplitude(self, t):
        return pulse((t + 0.5) * self.dt) if self.mode == "tm" else pulse(t * self.dt)

    def h_amplitude(self, t):
        if self.mode == "tm":
            return -pulse(t * self.dt - DX / 2)
        return pulse((t + 0.5) * self.dt - DX / 2)


def _eps(dtype):
    from meep_torch_amd import synth
    r = torch.tensor(synth.splitmix_uniform(NX * NY, 21).reshape(NX, NY), dtype=torch.float64)
    e = torch.ones(NX, NY, dtype=torch.float64)
    e[80:100, 20:44] = 1.0 + 3.0 * r[80:100, 20:44]      # a scatterer past the forward monitor
    return e.to(dtype)


def _hip(mode, dtype, eps, ck=0, n=N):
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE, FluxMonitorTM
    from meep_torch_amd.core.run import run_te, run_tm
    from meep_torch_amd.core.stepper_te import init_te_state
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dtype, device="cuda")
    src = Huygens(mode, g.dt, dtype, "cuda")
    if mode == "tm":
        m = build_tm_coeffs(g, eps_inf=eps.cuda(), npml=NPML)
        mons = [FluxMonitorTM(g, FREQS, x_normal=True, idx=i, lo=src.lo, hi=src.hi) for i in (75, 30)]
        st, acc = run_tm(init_tm_state(g), m, n, source=src, monitors=mons, checkpoint_every=ck)
    else:
        m = build_te_coeffs(g, eps_inf=eps.cuda(), npml=NPML)
        mons = [FluxMonitorTE(g, FREQS, ix=i, lo=src.lo, hi=src.hi) for i in (75, 30)]
        st, acc = run_te(init_te_state(g), m, n, source=src, monitors=mons, checkpoint_every=ck)
    return st, acc, [mon.flux(a, DX) for mon, a in zip(mons, acc)], src


def _oracle(mode, dtype, eps, n=N):
    from oracle import yee2d
    torch.set_num_threads(8)
    if mode == "tm":
        c = yee2d.tm_coeffs(NX, NY, DX, eps, NPML, dtype=dtype)
        st = yee2d.tm_state(NX, NY, dtype)
        kind, key = "tm_x", "hy_mask"
    else:
        c = yee2d.te_coeffs(NX, NY, DX, eps, NPML, dtype=dtype)
        st = yee2d.te_state(NX, NY, dtype)
        kind, key = "te_x", "hz_mask"
    src = Huygens(mode, c.dt, dtype, "cpu")
    mons = [{"type": "flux", "kind": kind, "idx": i, "lo": src.lo, "hi": src.hi, "freqs": FREQS} for i in (75, 30)]
    amps = [src.amplitude(t) for t in range(n)]
    hamps = [src.h_amplitude(t) for t in range(n)]
    s, acc = yee2d.run2d(mode, st, c, n, amps=amps, mask=src.mask, monitors=mons, hamps=hamps,
                         hmasks={key: next(iter(src.h_masks.values()))})
    return s, acc, [yee2d.flux_from_acc(kind, a, DX) for a in acc]


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("mode", ["tm", "te"])
@pytest.mark.parametrize("dtype,tf,ts", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)],
                         ids=["f64", "f32"])
def test_tfsf_run_matches_oracle(mode, dtype, tf, ts):
    eps = _eps(dtype)
    st, acc, flux, _ = _hip(mode, dtype, eps)
    s, oacc, oflux = _oracle(mode, dtype, eps)
    names = ("Ez", "Dz", "Hx", "Hy") if mode == "tm" else ("Hz", "Ex", "Ey", "Dx", "Dy")
    errs = {f: rel(getattr(st, f), s[f]) for f in names}
    errs.update({f"acc{i}": rel(a, b) for i, (a, b) in enumerate(zip(acc, oacc))})
    print(f"\n[tfsf {mode} {dtype}] " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for f in names:
        assert errs[f] <= tf, (f, errs[f])
    for i in range(2):
        assert errs[f"acc{i}"] <= ts and rel(flux[i], oflux[i]) <= ts


@pytest.mark.parametrize("mode", ["tm", "te"])
def test_tfsf_chunked_and_one_way(mode):
    """Chunks of 150 steps regroup only the monitor sums; in vacuum the pair
    radiates forward only (backward flux < 2e-3 of forward)."""
    e1 = torch.ones(NX, NY, dtype=torch.float64)
    st0, acc0, f0, _ = _hip(mode, torch.float64, e1, ck=0)
    st1, acc1, f1, _ = _hip(mode, torch.float64, e1, ck=150)
    for f in st0._fields:
        a, b = getattr(st0, f), getattr(st1, f)
        if a is not None and a.numel():
            assert torch.equal(a, b), f
    assert torch.allclose(f0[0], f1[0], rtol=1e-12, atol=0)
    fwd, bwd = f0[0].cpu(), f0[1].cpu()
    assert float(fwd[1]) > 0
    assert float((bwd.abs() / fwd).max()) < 2e-3


def test_step_wrappers_take_magnetic_masks():
    """step_tm / step_te with src_h_amp and the TFSF masks, one step from a
    random field state (psi zero: the engine keeps psi at zero outside the
    CPML shells, as every run does), against the oracle's step (FP64)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.stepper_te import FIELDS as TEF, init_te_state, step_te
    from meep_torch_amd.core.stepper_tm import FIELDS as TMF, init_tm_state, step_tm
    from oracle import yee2d
    dt = torch.float64
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dt, device="cuda")
    gen = torch.Generator().manual_seed(3)
    eps = _eps(dt)
    for mode in ("tm", "te"):
        src = Huygens(mode, g.dt, dt, "cuda")
        hm = next(iter(src.h_masks.values()))
        if mode == "tm":
            st = init_tm_state(g)
            st = st._replace(**{f: torch.randn(getattr(st, f).shape, generator=gen, dtype=dt).cuda()
                                for f in ("Ez", "Dz", "Hx", "Hy")})
            m = build_tm_coeffs(g, eps_inf=eps.cuda(), npml=NPML)
            out = step_tm(st, m, 0.7, src.mask, -0.3, src_hy_mask=hm)
            c = yee2d.tm_coeffs(NX, NY, DX, eps, NPML, dtype=dt)
            ref = yee2d.tm_step({f: getattr(st, f).cpu() for f in TMF}, c, c.dt * 0.7, src.mask.cpu(),
                                hcoef=c.dt * -0.3, hy_mask=hm.cpu())
            names = TMF
        else:
            st = init_te_state(g)
            st = st._replace(**{f: torch.randn(getattr(st, f).shape, generator=gen, dtype=dt).cuda()
                                for f in ("Hz", "Ex", "Ey", "Dx", "Dy")})
            m = build_te_coeffs(g, eps_inf=eps.cuda(), npml=NPML)
            out = step_te(st, m, 0.7, src.mask, 0.4, src_hz_mask=hm)
            c = yee2d.te_coeffs(NX, NY, DX, eps, NPML, dtype=dt)
            ref = yee2d.te_step({f: getattr(st, f).cpu() for f in TEF}, c, c.dt * 0.7, src.mask.cpu(),
                                hcoef=c.dt * 0.4, hz_mask=hm.cpu())
            names = TEF
        for f in names:
            assert rel(getattr(out, f), ref[f]) <= 1e-13, (mode, f)


@pytest.mark.parametrize("mode", ["tm", "te"])
def test_tfsf_eps_gradient_matches_oracle_autograd(mode):
    """The adjoint recompute replays the magnetic current: the eps gradient
    of the forward flux with a TFSF source equals CPU autograd through the
    oracle (FP64)."""
    from oracle import yee2d
    e0 = _eps(torch.float64)
    eps = e0.clone().cuda().requires_grad_(True)
    _, _, flux, _ = _hip(mode, torch.float64, eps, ck=100, n=300)
    flux[0].sum().backward()
    e_cpu = e0.clone().requires_grad_(True)
    _, _, oflux = _oracle(mode, torch.float64, e_cpu, n=300)
    g_ref, = torch.autograd.grad(oflux[0].sum(), e_cpu)
    assert float(g_ref.abs().max()) > 0
    e = rel(eps.grad, g_ref)
    print(f"\n[tfsf grad {mode}] rel-L2 {e:.2e}")
    assert e <= 1e-9


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batched_path_takes_magnetic_sources(dtype):
    """TFSF magnetic currents in the batched launch (per-step h-coefficient
    tables, the sheets injected by the tiled kernel): two variants, one with
    the Huygens pair and one without, equal run_tm of each alone bit for
    bit."""
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=NX, ny=NY, dx=DX, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=torch.ones(NX, NY, dtype=dtype, device="cuda"), npml=NPML)
    srcs = [Huygens("tm", g.dt, dtype, "cuda"), Huygens("tm", g.dt, dtype, "cuda")]
    srcs[1].h_masks = {}
    mk = lambda: [FluxMonitorTM(g, [0.8, 1.0, 1.2], x_normal=True, idx=90, lo=NPML, hi=NY - NPML)]
    n = 400
    bt = BatchedTM([(m, s, mk()) for s in srcs], n, checkpoint_every=0, adjoint=False)
    out = bt.forward([m.eps_inf, m.eps_inf])
    for s, o in zip(srcs, out):
        _, (ref,) = run_tm(init_tm_state(g), m, n, source=s, monitors=mk())
        assert torch.equal(o[0], ref), float((o[0] - ref).abs().max())
    assert not torch.equal(out[0][0], out[1][0])
