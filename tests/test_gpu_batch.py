"""Batched TM variants of uneven size (one persistent launch per block of
steps, HIP-graph replay, batched adjoint) against the single-problem engine:
spectra bit-identical to run_tm(checkpoint_every=K) per variant, graph replays
identical to fresh runs, batched eps gradients equal to the per-variant
adjoint's, and a central finite-difference check."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

SIZES = [(70, 60), (96, 90), (120, 64), (50, 52)]
N_STEPS, CK = 300, 100


def _problem(nx, ny, dtype, seed):
    from meep_torch_amd import synth
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=dtype, device="cuda")
    r = torch.tensor(synth.splitmix_uniform(nx * ny, seed).reshape(nx, ny), dtype=torch.float64)
    eps = (1.0 + 3.0 * r).to(dtype).cuda()
    m = build_tm_coeffs(g, eps_inf=eps, npml=8)
    src = LineSource(g, ix=12, j0=10, j1=ny - 10, profile=GaussianPulse(1.0, 0.5, cutoff=2.0))
    mon = FluxMonitorTM(g, [0.8, 1.0, 1.2], x_normal=True, idx=nx - 14, lo=9, hi=ny - 9)
    return g, m, src, mon, eps


def _single(g, m, src, mon, eps):
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    _, (acc,) = run_tm(init_tm_state(g), m._replace(eps_inf=eps), N_STEPS, source=src, monitors=[mon],
                       checkpoint_every=CK)
    return acc


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batched_forward_matches_run_tm(dtype):
    from meep_torch_amd.batch import BatchedTM
    probs = [_problem(nx, ny, dtype, 10 + i) for i, (nx, ny) in enumerate(SIZES)]
    bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    out = bt.forward([p[4] for p in probs])
    for (g, m, src, mon, eps), o in zip(probs, out):
        ref = _single(g, m, src, mon, eps)
        assert torch.equal(o[0], ref), float((o[0] - ref).abs().max())


def test_graph_replay_matches_fresh_runs():
    from meep_torch_amd.batch import BatchedTM
    dtype = torch.float32
    probs = [_problem(nx, ny, dtype, 20 + i) for i, (nx, ny) in enumerate(SIZES)]
    bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    bt.set_eps([p[4] for p in probs])
    bt.capture()
    for it in range(3):
        eps_list = [p[4] * (1.0 + 0.1 * it) for p in probs]
        out = bt.forward(eps_list)
        for (g, m, src, mon, _), e, o in zip(probs, eps_list, out):
            ref = _single(g, m, src, mon, e)
            assert torch.equal(o[0], ref), (it, float((o[0] - ref).abs().max()))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_gradient_matches_single_adjoint(dtype):
    from meep_torch_amd.batch import BatchedTM
    probs = [_problem(nx, ny, dtype, 30 + i) for i, (nx, ny) in enumerate(SIZES)]
    bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    eps_b = [p[4].clone().requires_grad_(True) for p in probs]
    out = bt.apply(eps_b)
    J = sum(mon.flux(o[0], 0.05).sum() for (_, _, _, mon, _), o in zip(probs, out))
    J.backward()
    for (g, m, src, mon, eps), eb in zip(probs, eps_b):
        e1 = eps.clone().requires_grad_(True)
        acc = _single(g, m, src, mon, e1)
        mon.flux(acc, 0.05).sum().backward()
        rel = float((eb.grad - e1.grad).norm() / e1.grad.norm())
        assert rel <= (1e-12 if dtype == torch.float64 else 1e-6), rel


def test_batched_gradient_central_fd():
    from meep_torch_amd.batch import BatchedTM
    dtype = torch.float64
    probs = [_problem(nx, ny, dtype, 40 + i) for i, (nx, ny) in enumerate(SIZES[:2])]
    bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    eps_b = [p[4].clone().requires_grad_(True) for p in probs]
    out = bt.apply(eps_b)
    (mon1 := probs[1][3])
    J = mon1.flux(out[1][0], 0.05).sum()
    J.backward()
    h = 1e-4
    for (i, j) in ((40, 45), (55, 30), (70, 60)):
        ep = [p[4].clone() for p in probs]
        em = [p[4].clone() for p in probs]
        ep[1][i, j] += h
        em[1][i, j] -= h
        fp = float(mon1.flux(bt.forward(ep)[1][0], 0.05).sum())
        fm = float(mon1.flux(bt.forward(em)[1][0], 0.05).sum())
        fd = (fp - fm) / (2 * h)
        g = float(eps_b[1].grad[i, j])
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-12) + 1e-14, ((i, j), g, fd)
    assert float(eps_b[0].grad.abs().max()) == 0.0


BIG = [(336, 310), (392, 448)]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batched_interior_tiles_match_two_pass(dtype, monkeypatch):
    """Grids large enough that the batched launch's interior-tile kernel
    (tm_batch<..., PML = false>, the FP32 32 x 64 x 4-step tile the C4 design
    loop runs) carries about half the cells or more: spectra bit-identical to the two-pass
    per-half-step path, and the batched gradient equal to the single-problem
    adjoint's."""
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.monitors import FluxMonitorTM
    probs = []
    for i, (nx, ny) in enumerate(BIG):
        g, m, src, _, eps = _problem(nx, ny, dtype, 50 + i)
        # 300 steps reach about 90 cells: the monitor sits where the pulse
        # arrives, past the first rows of interior tiles
        mon = FluxMonitorTM(g, [0.8, 1.0, 1.2], x_normal=True, idx=80, lo=9, hi=ny - 9)
        probs.append((g, m, src, mon, eps))
    bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    for (g, m, src, mon, eps), (nx, ny) in zip(probs, BIG):
        n_in, n_ring = engine2d.tile_counts("tm", nx, ny, m, dtype, bt.variant)
        assert n_in >= 25 and n_ring > 0, (n_in, n_ring)   # interior tiles: about half the grid or more
    eps_b = [p[4].clone().requires_grad_(True) for p in probs]
    out = bt.apply(eps_b)
    J = sum(mon.flux(o[0], 0.05).sum() for (_, _, _, mon, _), o in zip(probs, out))
    J.backward()
    monkeypatch.setenv("MEEP_TORCH_AMD_VARIANT", "-1")
    for (g, m, src, mon, eps), o, eb in zip(probs, out, eps_b):
        e1 = eps.clone().requires_grad_(True)
        ref = _single(g, m, src, mon, e1)
        assert torch.equal(o[0].detach(), ref.detach()), float((o[0] - ref).abs().max())
        assert float(ref.detach().abs().max()) > 0
        mon.flux(ref, 0.05).sum().backward()
        rel = float((eb.grad - e1.grad).norm() / e1.grad.norm())
        assert rel <= (1e-12 if dtype == torch.float64 else 1e-6), rel


def test_adjoint_graph_replays_match_eager():
    """The reverse sweep captured in a HIP graph (descriptors uploaded once,
    launches only): three design iterations with new eps and new cotangents,
    eager work and synchronisations between the replays; every replayed
    gradient equals the eager gradient of the same iteration bit for bit,
    and the forward graph's spectra equal fresh runs."""
    from meep_torch_amd.batch import BatchedTM
    dtype = torch.float32
    probs = [_problem(nx, ny, dtype, 60 + i) for i, (nx, ny) in enumerate(SIZES)]
    mk = lambda: BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=CK)
    eager, graphed = mk(), mk()
    graphed.set_eps([p[4] for p in probs])
    graphed.capture()
    graphed.capture_adjoint()
    for it in range(3):
        eps_list = [p[4] * (1.0 + 0.05 * it) for p in probs]
        res = []
        for bt in (eager, graphed):
            eb = [e.clone().requires_grad_(True) for e in eps_list]
            out = bt.apply(eb)
            J = sum((1.0 + 0.3 * it + 0.1 * v) * mon.flux(o[0], 0.05).sum()
                    for v, ((_, _, _, mon, _), o) in enumerate(zip(probs, out)))
            J.backward()
            torch.cuda.synchronize()
            res.append(([o[0].detach().clone() for o in out], [e.grad.clone() for e in eb]))
        (oe, ge), (og, gg) = res
        for a, b in zip(oe, og):
            assert torch.equal(a, b), it
        for a, b in zip(ge, gg):
            assert torch.equal(a, b), (it, float((a - b).abs().max()))
        assert float(ge[0].abs().max()) > 0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("ck", [CK, 75])
def test_fused_adjoint_step_equals_three_passes(dtype, ck):
    """The fused reverse step (one tiled pass per step, two cotangent sets,
    transposed differences in LDS) against the three transpose passes: eps
    gradients bit for bit, with segments of an even (100) and an odd (75,
    the result copied back from the second set) number of steps, on grids
    that end in partial tiles in both directions."""
    from meep_torch_amd.batch import BatchedTM
    sizes = SIZES + [(83, 131)]
    probs = [_problem(nx, ny, dtype, 60 + i) for i, (nx, ny) in enumerate(sizes)]
    grads = []
    for fused in (False, True):
        bt = BatchedTM([(m, src, [mon]) for (g, m, src, mon, eps) in probs], N_STEPS, checkpoint_every=ck,
                       fused_adjoint=fused)
        eps_b = [p[4].clone().requires_grad_(True) for p in probs]
        out = bt.apply(eps_b)
        J = sum(mon.flux(o[0], 0.05).sum() for (_, _, _, mon, _), o in zip(probs, out))
        J.backward()
        assert bt._adj["fused"] == fused
        grads.append([e.grad.clone() for e in eps_b])
    for a, b in zip(*grads):
        assert float(a.abs().max()) > 0
        assert torch.equal(a, b), float((a - b).abs().max())
