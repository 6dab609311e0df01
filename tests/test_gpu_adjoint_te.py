"""Gradients of the hand-written TE adjoint (mt_te2d_adjoint behind
torch.autograd.Function, reached through run_te when eps requires grad).

Setting: the 64 x 48 TE case of the reference's golden vectors (8-cell CPML,
line source on Dy, 400 steps).  The oracle's TE step is pinned bit-exactly to
those golden vectors; its gradient is taken by torch autograd on the CPU, one
objective at a time, so each adjoint path is checked on its own: the x- and
y-normal flux monitors (E term at half steps, averaged Hz term), DFT field
monitors of Hz and of Ey, and a weight on the final Hz (state cotangent).
Bars: FP64 rel-L2 <= 1e-9 against the oracle's FP64 gradient; FP32 rel-L2
<= 1e-4 against the oracle's FP32 gradient and <= 1e-3 against FP64; central
finite differences of the HIP forward (FP64, h = 1e-4) within 1e-6."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

FREQS5 = [0.3, 0.4, 0.5, 0.6, 0.7]
OBJECTIVES = ("flux_x", "flux_y", "dft_hz", "dft_ey", "state_hz")


def _weights(dtype):
    g = torch.Generator().manual_seed(5)
    return torch.randn(63, 47, generator=g, dtype=torch.float64).to(dtype)


def _objective(name, flux_x, flux_y, acc_hz, acc_ey, hz, dtype):
    if name == "flux_x":
        return flux_x.sum()
    if name == "flux_y":
        return flux_y.sum()
    if name == "dft_hz":
        return (acc_hz ** 2).sum() * 1e2
    if name == "dft_ey":
        return (acc_ey ** 2).sum() * 1e2
    return (hz * _weights(dtype).to(hz.device)).sum()


def _oracle_grads(dtype):
    """CPU autograd through the oracle's TE run, one gradient per objective."""
    from oracle import yee2d
    G = golden("te_small_f64")
    torch.set_num_threads(8)
    eps = torch.from_numpy(np.ascontiguousarray(G["eps"])).to(dtype).requires_grad_(True)
    c = yee2d.te_coeffs(64, 48, 0.05, eps, 8, dtype=dtype)
    mons = [{"type": "flux", "kind": "te_x", "idx": 50, "lo": 10, "hi": 38, "freqs": FREQS5},
            {"type": "flux", "kind": "te_y", "idx": 36, "lo": 10, "hi": 54, "freqs": FREQS5},
            {"type": "dft", "comp": "Hz", "freqs": [0.45, 0.55]},
            {"type": "dft", "comp": "Ey", "freqs": [0.5]}]
    mask = torch.from_numpy(np.ascontiguousarray(G["mask"])).to(dtype)
    s, acc = yee2d.run2d("te", yee2d.te_state(64, 48, dtype), c, 400, amps=G["amp"], mask=mask,
                         monitors=mons)
    fx = yee2d.flux_from_acc("te_x", acc[0], 0.05)
    fy = yee2d.flux_from_acc("te_y", acc[1], 0.05)
    out = {}
    for name in OBJECTIVES:
        J = _objective(name, fx, fy, acc[2], acc[3], s["Hz"], dtype)
        out[name] = torch.autograd.grad(J, eps, retain_graph=True)[0].double()
    return out


_CACHE = {}


def oracle_grads(dtype):
    if dtype not in _CACHE:
        _CACHE[dtype] = _oracle_grads(dtype)
    return _CACHE[dtype]


def _hip_run(eps, dtype, ck=0):
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    g = Grid(nx=64, ny=48, dx=0.05, dtype=dtype, device="cuda")
    m = build_te_coeffs(g, eps_inf=eps, npml=8)
    src = LineSource(g, ix=12, j0=16, j1=32, profile=GaussianPulse(0.5, 1.0))
    src.mask = src.mask[:, :-1].contiguous()
    mons = [FluxMonitorTE(g, FREQS5, ix=50, lo=10, hi=38),
            FluxMonitorTEY(g, FREQS5, iy=36, lo=10, hi=54),
            DFTFieldsMonitor(g, [0.45, 0.55], component="Hz"),
            DFTFieldsMonitor(g, [0.5], component="Ey")]
    st, acc = run_te(init_te_state(g), m, 400, source=src, monitors=mons, checkpoint_every=ck)
    return mons[0].flux(acc[0], g.dx), mons[1].flux(acc[1], g.dx), acc[2], acc[3], st.Hz


def _hip_grad(name, dtype, ck=0):
    G = golden("te_small_f64")
    eps = torch.from_numpy(np.ascontiguousarray(G["eps"])).to(dtype).cuda().requires_grad_(True)
    J = _objective(name, *_hip_run(eps, dtype, ck), dtype)
    J.backward()
    return eps.grad.detach().double().cpu()


def rel(a, b):
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("name", OBJECTIVES)
def test_te_grad_matches_oracle_autograd_f64(name):
    ref = oracle_grads(torch.float64)[name]
    got = _hip_grad(name, torch.float64)
    assert float(ref.abs().max()) > 0
    e = rel(got, ref)
    print(f"\n[te adjoint f64] {name}: rel-L2 {e:.2e}")
    assert e <= 1e-9, (name, e)


@pytest.mark.parametrize("name", ("flux_x", "dft_hz", "state_hz"))
def test_te_grad_matches_oracle_autograd_f32(name):
    got = _hip_grad(name, torch.float32)
    e32 = rel(got, oracle_grads(torch.float32)[name])
    e64 = rel(got, oracle_grads(torch.float64)[name])
    print(f"\n[te adjoint f32] {name}: vs oracle f32 {e32:.2e}, vs oracle f64 {e64:.2e}")
    assert e32 <= 1e-4 and e64 <= 1e-3, (name, e32, e64)


def test_te_grad_checkpoint_chunks_agree():
    """checkpoint_every regroups the monitor sums only: gradients agree."""
    a = _hip_grad("flux_x", torch.float64, ck=0)
    b = _hip_grad("flux_x", torch.float64, ck=150)
    assert rel(b, a) <= 1e-11


def test_te_grad_matches_central_fd():
    """FP64, h = 1e-4 on four cells around the source line and the monitors;
    J = flux_x + flux_y (both E/H flux paths at once)."""
    G = golden("te_small_f64")
    eps0 = torch.from_numpy(np.ascontiguousarray(G["eps"])).cuda()

    def J(e):
        fx, fy, *_ = _hip_run(e, torch.float64)
        return fx.sum() + fy.sum()

    eps = eps0.clone().requires_grad_(True)
    J(eps).backward()
    h = 1e-4
    for (i, j) in ((30, 24), (45, 30), (20, 20), (50, 40)):
        ep, em = eps0.clone(), eps0.clone()
        ep[i, j] += h
        em[i, j] -= h
        with torch.no_grad():
            fd = (float(J(ep)) - float(J(em))) / (2 * h)
        g = float(eps.grad[i, j])
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-12) + 1e-14, ((i, j), g, fd)


def test_te_grad_only_eps_supported():
    """A grad-requiring PML profile is refused loudly (no silent zero gradient)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    g = Grid(nx=32, ny=32, dx=0.05, dtype=torch.float64, device="cuda")
    m = build_te_coeffs(g, eps_inf=2.0, npml=4)
    m = m._replace(b_hx=m.b_hx.clone().requires_grad_(True))
    with pytest.raises(NotImplementedError):
        run_te(init_te_state(g), m, 10)
