"""3D hand-written adjoint (run_3d_adjoint on mt_yee3d_adjoint) against the
reference's checkpointed autograd gradient (golden grad_3d_f64: 30x24x20,
6-cell CPML, point Ez source, 90 steps in chunks of 16, one-frequency Ez
plane DFT, L = sum |E|^2), central finite differences in float64, chunk
invariance, FP32 against FP64, and gradients routed into an upstream graph."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def _setup(eps, dtype):
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, PointSource3D
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    nx, ny, nz = 30, 24, 20
    c = build_3d_coeffs(nx, ny, nz, 0.05, eps=eps, npml=6, dtype=dtype, device="cuda")
    src = PointSource3D(c.dt, "Ez", nx // 2, ny // 2, nz // 2 - 3, GaussianPulse(0.6, 0.25))
    mon = DFTFieldsMonitor3D(c.dt, [0.6], "Ez", plane=(0, nx - 10), dtype=dtype)
    st0 = init_3d_state(nx, ny, nz, npml=6, dtype=dtype, device="cuda")
    return c, src, mon, st0


def _loss(accs):
    return (accs[0][0, 0] ** 2 + accs[0][0, 1] ** 2).sum()


def _run(eps, dtype, chunk=16, steps=90):
    from meep_torch_amd.core.adjoint3d import run_3d_adjoint
    c, src, mon, st0 = _setup(eps, dtype)
    return run_3d_adjoint(c, steps, chunk, src, [mon], _loss, st0)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_matches_reference_autograd_f64():
    g = golden("grad_3d_f64")
    eps = torch.tensor(g["eps"]).cuda().requires_grad_(True)
    L, accs = _run(eps, torch.float64)
    assert abs(L - float(g["L"])) <= 1e-12 * abs(float(g["L"]))
    assert _rel(accs[0].cpu(), g["acc"]) <= 1e-12
    e = _rel(eps.grad.cpu(), g["grad"])
    print(f"\n[adjoint3d f64] grad rel-L2 vs reference autograd {e:.2e}")
    assert e <= 1e-10


def test_central_fd_f64():
    g = golden("grad_3d_f64")
    base = torch.tensor(g["eps"]).cuda()
    eps = base.clone().requires_grad_(True)
    _run(eps, torch.float64)
    grad = eps.grad.cpu().numpy()
    h = 1e-4
    for idx in [(15, 12, 8), (14, 11, 10), (20, 12, 7), (9, 9, 9)]:
        ep, em = base.clone(), base.clone()
        ep[idx] += h
        em[idx] -= h
        fd = (_run(ep, torch.float64)[0] - _run(em, torch.float64)[0]) / (2 * h)
        assert abs(grad[idx] - fd) <= 1e-6 * max(abs(fd), 1e-30) + 1e-14 * np.abs(grad).max(), (idx, grad[idx], fd)


def test_chunk_invariance_and_fp32():
    g = golden("grad_3d_f64")
    e1 = torch.tensor(g["eps"]).cuda().requires_grad_(True)
    _run(e1, torch.float64, chunk=7)
    assert _rel(e1.grad.cpu(), g["grad"]) <= 1e-10
    e32 = torch.tensor(g["eps"]).float().cuda().requires_grad_(True)
    _run(e32, torch.float32)
    e = _rel(e32.grad.cpu(), g["grad"])
    print(f"\n[adjoint3d f32] grad rel-L2 vs f64 reference {e:.2e}")
    assert e <= 1e-4


def test_gradient_reaches_upstream_graph():
    g = golden("grad_3d_f64")
    w = torch.tensor(2.0, dtype=torch.float64, device="cuda", requires_grad=True)
    base = torch.tensor(g["eps"]).cuda()
    eps = base * (w / 2.0)
    _run(eps, torch.float64)
    ref = (torch.tensor(g["grad"]).cuda() * base / 2.0).sum()
    assert abs(float(w.grad) - float(ref)) <= 1e-10 * abs(float(ref))


@pytest.mark.parametrize("poles", [False, True], ids=["sigma", "sigma+poles"])
def test_media_gradient_matches_oracle_autograd(poles):
    """With sigma_d and sigma_b (E' = d1 E + (d2 curl + source) / eps_c,
    H' = h1 H - h2 curl), and with a Lorentz and a Drude pole on top (the
    pole stacks' cotangents carried between steps), the 3D adjoint's eps
    gradient equals CPU autograd through the oracle's 3D steps (FP64),
    across checkpoint chunks."""
    from meep_torch_amd.core.adjoint3d import run_3d_adjoint
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, PointSource3D
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    from oracle import yee3d
    nx, ny, nz, steps = 30, 24, 20, 90
    gen = torch.Generator().manual_seed(8)
    inner = torch.zeros(nx, ny, nz, dtype=torch.float64)
    inner[9:21, 8:16, 7:13] = 1.0
    eps0 = 1.0 + 2.0 * torch.rand(nx, ny, nz, generator=gen, dtype=torch.float64) * inner
    sd = 2.0 * torch.rand(nx, ny, nz, generator=gen, dtype=torch.float64) * inner
    sb = 1.0 * torch.rand(nx, ny, nz, generator=gen, dtype=torch.float64) * inner
    half = (torch.rand(nx, ny, nz, generator=gen, dtype=torch.float64) > 0.5).double() * inner
    pk = dict(poles=[(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)], pole_masks=[half, inner - half]) if poles else {}
    eps = eps0.clone().cuda().requires_grad_(True)
    c = build_3d_coeffs(nx, ny, nz, 0.05, eps=eps, npml=6, sigma_d=sd.cuda(), sigma_b=sb.cuda(),
                        dtype=torch.float64, device="cuda",
                        **({k: ([m.cuda() for m in v] if k == "pole_masks" else v) for k, v in pk.items()}))
    prof = GaussianPulse(0.6, 0.25)
    src = PointSource3D(c.dt, "Ez", nx // 2, ny // 2, nz // 2 - 3, prof)
    mon = DFTFieldsMonitor3D(c.dt, [0.6], "Ez", plane=(0, nx - 10), dtype=torch.float64)
    st0 = init_3d_state(nx, ny, nz, npml=6, n_poles=2 if poles else 0, dtype=torch.float64, device="cuda")
    run_3d_adjoint(c, steps, 16, src, [mon], _loss, st0)
    torch.set_num_threads(8)
    e = eps0.clone().requires_grad_(True)
    oc = yee3d.coeffs3d(nx, ny, nz, 0.05, e, 6, torch.float64, sigma_d=sd, sigma_b=sb, **pk)
    amps = [prof.amplitude((t + 0.5) * oc["dt"]) for t in range(steps)]
    _, oacc = yee3d.run3d(yee3d.state3d(nx, ny, nz, torch.float64, n_poles=2 if poles else 0), oc, steps, amps,
                          ("Ez", nx // 2, ny // 2, nz // 2 - 3),
                          [{"type": "dft", "comp": "Ez", "plane": (0, nx - 10), "freq": 0.6}])
    g_ref, = torch.autograd.grad(_loss(oacc), e)
    err = _rel(eps.grad.cpu(), g_ref)
    print(f"\n[adjoint3d media poles={poles}] grad rel-L2 vs oracle autograd {err:.2e}")
    assert float(g_ref[9:21, 8:16, 7:13].abs().max()) > 0 and err <= 1e-9
