"""The temporally blocked 2D steppers (every tile variant) must reproduce the
two-pass per-half-step path bit for bit: fields, psi, monitor accumulators,
probe stacks and the adjoint tape, on ragged grids (sizes not multiples of the
tile), with CPML shells, sources inside halo zones and step counts that are
not multiples of the steps-per-launch."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NV = 7


def _eps(nx, ny, dtype):
    x = torch.arange(nx, dtype=torch.float64)[:, None]
    y = torch.arange(ny, dtype=torch.float64)[None, :]
    e = 1.0 + 2.0 * torch.exp(-((x - nx / 2) ** 2 + (y - ny / 3) ** 2) / 200.0)
    e[:, ny // 2:ny // 2 + 6] = 12.0
    return e.to(dtype).cuda()


def _run(mode, dtype, variant, k, y_pmc=False):
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import (DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY,
                                              FluxMonitorTM, PointProbe)
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import FIELDS as TEF, init_te_state
    from meep_torch_amd.core.stepper_tm import FIELDS as TMF, init_tm_state
    nx, ny = 151, 133
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=dtype, device="cuda")
    fr = [0.6, 0.8, 1.0]
    src = LineSource(g, ix=20, j0=17, j1=90, profile=GaussianPulse(0.8, 0.5, cutoff=2.0))
    if mode == "tm":
        m = build_tm_coeffs(g, eps_inf=_eps(nx, ny, dtype), npml=(10, 0) if y_pmc else 10, y_pmc=y_pmc)
        st = init_tm_state(g, y_pmc=y_pmc)
        names = TMF
        mons = [FluxMonitorTM(g, fr, x_normal=True, idx=63, lo=5, hi=120),
                FluxMonitorTM(g, fr, x_normal=False, idx=64, lo=3, hi=149),
                DFTFieldsMonitor(g, [0.9], component="Ez"), DFTFieldsMonitor(g, [0.9], component="Hy"),
                PointProbe(70, 40)]
    else:
        m = build_te_coeffs(g, eps_inf=_eps(nx, ny, dtype), npml=10)
        st = init_te_state(g)
        src.mask = src.mask[:, :-1].contiguous()
        names = TEF
        mons = [FluxMonitorTE(g, fr, ix=63, lo=5, hi=120), FluxMonitorTEY(g, fr, iy=64, lo=3, hi=149),
                DFTFieldsMonitor(g, [0.9], component="Hz"), DFTFieldsMonitor(g, [0.9], component="Ey")]
    work = {f: getattr(st, f).clone() for f in names}
    plan = engine2d.source_plan(src, work["Dz" if mode == "tm" else "Dy"].shape, "cuda", dtype)
    tape = torch.zeros(k, nx, ny, dtype=dtype, device="cuda") if mode == "tm" else None
    outs = engine2d.run_chunk(mode, work, m, 0, k, plan, mons, tape=tape, variant=variant)
    return work, outs, tape


@pytest.mark.parametrize("variant", range(NV))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("mode", ["tm", "te"])
def test_tiled_matches_two_pass(mode, dtype, variant):
    k = 77
    ref_w, ref_o, ref_t = _run(mode, dtype, -1, k)
    w, o, t = _run(mode, dtype, variant, k)
    for f in ref_w:
        assert torch.equal(w[f], ref_w[f]), (f, float((w[f] - ref_w[f]).abs().max()))
    for a, b in zip(o, ref_o):
        assert torch.equal(a, b), float((a - b).abs().max())
    if ref_t is not None:
        assert torch.equal(t, ref_t)


@pytest.mark.parametrize("variant", [0, 2, 5])
def test_tiled_tm_y_pmc(variant):
    ref_w, ref_o, _ = _run("tm", torch.float32, -1, 41, y_pmc=True)
    w, o, _ = _run("tm", torch.float32, variant, 41, y_pmc=True)
    for f in ref_w:
        assert torch.equal(w[f], ref_w[f]), f
    for a, b in zip(o, ref_o):
        assert torch.equal(a, b)


def test_autotuner_picks_a_valid_variant():
    from meep_torch_amd import tune
    from meep_torch_amd.core import engine2d
    w, _, _ = _run("te", torch.float32, -1, 1)
    v = tune.variant_for("te", w, None, None, [], 10) if False else None
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    g = Grid(nx=300, ny=280, dx=0.05, dtype=torch.float32, device="cuda")
    m = build_te_coeffs(g, eps_inf=2.0, npml=12)
    from meep_torch_amd.core.stepper_te import FIELDS, init_te_state
    st = init_te_state(g)
    work = {f: getattr(st, f).clone() for f in FIELDS}
    best = tune.tune("te", work, m, None, [])
    assert -1 <= best < len(tune.variants(torch.float32))
    assert set(tune.tune.last) == set(range(-1, len(tune.variants(torch.float32))))
