"""Gradients of the hand-written TM adjoint (mt_tm2d_adjoint behind
torch.autograd.Function) against the reference's own autograd gradient
(golden vectors) and central finite differences.  Bars: FP64 rel-L2 <= 1e-9
vs the reference; FP32 rel-L2 <= 1e-4 vs the reference's FP32 gradient and
<= 1e-3 vs its FP64 gradient; FD agreement <= 1e-6 (FP64)."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def rel(a, b):
    a = a.detach().double().cpu()
    b = torch.as_tensor(np.asarray(b)).double()
    return float((a - b).norm() / b.norm())


def _flux_objective(eps, dtype, ck=200):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=48, ny=48, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=eps, npml=6)
    src = LineSource(g, ix=8, j0=18, j1=30, profile=GaussianPulse(1.0, 0.4), amp_func=lambda y: 5.0)
    mon = FluxMonitorTM(g, [0.9, 1.0, 1.1], x_normal=True, idx=34, lo=16, hi=32)
    _, (acc,) = run_tm(init_tm_state(g), m, 1400, source=src, monitors=[mon], checkpoint_every=ck)
    return mon.flux(acc, g.dx).sum(), acc


@pytest.mark.parametrize("tag,dtype", [("f64", torch.float64), ("f32", torch.float32)])
def test_grad_matches_reference_autograd(tag, dtype):
    G = golden(f"grad_tm_{tag}")
    eps = torch.from_numpy(G["eps"]).cuda().requires_grad_(True)
    J, acc = _flux_objective(eps, dtype)
    J.backward()
    e_acc = rel(acc, G["acc"])
    e_g = rel(eps.grad, G["grad"])
    G64 = golden("grad_tm_f64")
    e_g64 = rel(eps.grad, G64["grad"])
    print(f"grad {tag}: acc rel {e_acc:.2e}, grad vs ref-{tag} {e_g:.2e}, vs ref-f64 {e_g64:.2e}")
    if dtype == torch.float64:
        assert e_acc <= 1e-10 and e_g <= 1e-9
    else:
        assert e_acc <= 1e-4 and e_g <= 1e-4 and e_g64 <= 1e-3


def test_grad_matches_central_fd():
    """FP64, h = 1e-4, four pixels inside/around the scatterer."""
    G = golden("grad_tm_f64")
    eps0 = torch.from_numpy(G["eps"]).cuda()
    eps = eps0.clone().requires_grad_(True)
    J, _ = _flux_objective(eps, torch.float64)
    J.backward()
    h = 1e-4
    for (i, j) in ((24, 24), (20, 27), (30, 22), (17, 24)):
        ep, em = eps0.clone(), eps0.clone()
        ep[i, j] += h
        em[i, j] -= h
        with torch.no_grad():
            fd = (float(_flux_objective(ep, torch.float64)[0]) - float(_flux_objective(em, torch.float64)[0])) / (2 * h)
        g = float(eps.grad[i, j])
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-12) + 1e-14, ((i, j), g, fd)


def test_grad_without_checkpoint_chunks_equals_chunked():
    """checkpoint_every only changes summation grouping: gradients agree."""
    G = golden("grad_tm_f64")
    e1 = torch.from_numpy(G["eps"]).cuda().requires_grad_(True)
    e2 = torch.from_numpy(G["eps"]).cuda().requires_grad_(True)
    _flux_objective(e1, torch.float64, ck=0)[0].backward()
    _flux_objective(e2, torch.float64, ck=350)[0].backward()
    assert rel(e1.grad, e2.grad.cpu().numpy()) <= 1e-11


def test_design_chain_gradient():
    """rho -> conic filter -> tanh projection -> paint -> run_tm -> flux:
    autograd reaches rho through the adjoint; checked by central FD."""
    from meep_torch_amd.adjoint.filters import conic_filter, tanh_projection
    from meep_torch_amd import synth
    rho0 = torch.tensor(synth.splitmix_uniform(12 * 12, 3).reshape(12, 12), dtype=torch.float64).cuda()
    base = torch.ones(48, 48, dtype=torch.float64, device="cuda")

    def J(rho):
        r = tanh_projection(conic_filter(rho, 0.1, 0.05), 4.0, 0.5)
        eps = base.clone()
        eps[18:30, 18:30] = 1.0 + 3.0 * r
        return _flux_objective(eps, torch.float64)[0]

    rho = rho0.clone().requires_grad_(True)
    J(rho).backward()
    h = 1e-5
    for (i, j) in ((5, 5), (2, 9)):
        rp, rm = rho0.clone(), rho0.clone()
        rp[i, j] += h
        rm[i, j] -= h
        with torch.no_grad():
            fd = (float(J(rp)) - float(J(rm))) / (2 * h)
        g = float(rho.grad[i, j])
        assert abs(g - fd) <= 1e-5 * max(abs(fd), 1e-12) + 1e-14, ((i, j), g, fd)
