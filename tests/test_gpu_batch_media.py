"""Batched TM runs whose variants carry media (ADE poles, chi3, sigma_d,
sigma_b / mu): BatchedTM runs them one after the other through the
single-problem path, so its spectra equal run_tm's per variant and its eps
gradients equal autograd through run_tm (the hand-written TM adjoint) bit
for bit; apply() gives the same gradients through backward()."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, N = 64, 48, 300


def _problem(v, dtype):
    from meep_torch_amd import synth
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    g = Grid(nx=NX, ny=NY, dx=0.05, dtype=dtype, device="cuda")
    r = torch.tensor(synth.splitmix_uniform(NX * NY, 50 + v).reshape(NX, NY), dtype=torch.float64)
    blk = torch.zeros(NX, NY, dtype=torch.float64)
    blk[24:40, 12:36] = 1.0
    eps = (1.0 + 2.0 * blk * r).to(dtype).cuda()
    m = build_tm_coeffs(g, eps_inf=eps, npml=8, poles=[(1.2, 0.1, 1.5)], pole_masks=[(blk * (r > 0.4)).to(dtype).cuda()],
                        chi3=(0.02 * blk).to(dtype).cuda(), sigma_d=(0.2 * blk).to(dtype).cuda(),
                        sigma_b=(0.1 * blk).to(dtype).cuda(), mu=(1.0 + 0.3 * blk).to(dtype).cuda())
    src = LineSource(g, ix=12, j0=9, j1=NY - 9, profile=GaussianPulse(1.0, 0.6))
    mon = FluxMonitorTM(g, [0.8, 1.0, 1.2], x_normal=True, idx=50, lo=9, hi=NY - 9)
    return g, m, src, mon, eps


def _reference(g, m, src, mon, eps, K):
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    e = eps.clone().requires_grad_(True)
    _, (acc,) = run_tm(init_tm_state(g, n_poles=1), m._replace(eps_inf=e), N, source=src, monitors=[mon],
                       checkpoint_every=K)
    w = torch.linspace(0.5, 1.5, acc.numel(), dtype=acc.dtype, device=acc.device).reshape(acc.shape)
    (ge,) = torch.autograd.grad((acc * w).sum(), [e])
    return acc.detach(), ge, w


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_media_batch_equals_single_runs(dtype):
    from meep_torch_amd.batch import BatchedTM
    K = 100
    probs = [_problem(v, dtype) for v in range(3)]
    bt = BatchedTM([(m, src, [mon]) for (_, m, src, mon, _) in probs], N, checkpoint_every=K)
    out = bt.forward([p[4] for p in probs])
    refs = [_reference(*p, K) for p in probs]
    for (acc,), (racc, _, _) in zip(out, refs):
        assert torch.equal(acc, racc)
        assert float(racc.abs().max()) > 0
    grads = bt.gradient([[w] for (_, _, w) in refs])
    for ge, (_, rge, _) in zip(grads, refs):
        assert torch.equal(ge, rge)
        assert float(rge.abs().max()) > 0
    # apply(): the same through backward()
    es = [p[4].clone().requires_grad_(True) for p in probs]
    res = bt.apply(es)
    loss = sum((a[0] * w).sum() for a, (_, _, w) in zip(res, refs))
    loss.backward()
    for e, (_, rge, _) in zip(es, refs):
        assert torch.equal(e.grad, rge)
