"""The temporally blocked 2D steppers (every tile variant) must reproduce the
two-pass per-half-step path bit for bit: fields, psi, monitor accumulators,
probe stacks and the adjoint tape, on ragged grids (sizes not multiples of the
tile), with CPML shells, sources inside halo zones and step counts that are
not multiples of the steps-per-launch.  The 151 x 133 grid runs ring tiles
only; the 320 x 300 grid gives every variant interior tiles (the psi-free
kernels the large configurations spend their time in), which each case
asserts through the library's tile-count query."""
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


def _run(mode, dtype, variant, k, y_pmc=False, size=(151, 133), expect_interior=None):
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import (DFTFieldsMonitor, FluxMonitorTE, FluxMonitorTEY,
                                              FluxMonitorTM, PointProbe)
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import FIELDS as TEF, init_te_state
    from meep_torch_amd.core.stepper_tm import FIELDS as TMF, init_tm_state
    nx, ny = size
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
    if expect_interior is not None and variant >= 0:
        n_in, n_ring = engine2d.tile_counts(mode, nx, ny, m, dtype, variant)
        assert (n_in > 0) == expect_interior, (n_in, n_ring)
    work = {f: getattr(st, f).clone() for f in names}
    plan = engine2d.source_plan(src, work["Dz" if mode == "tm" else "Dy"].shape, "cuda", dtype)
    tape = torch.zeros(k, nx, ny, dtype=dtype, device="cuda") if mode == "tm" else None
    outs = engine2d.run_chunk(mode, work, m, 0, k, plan, mons, tape=tape, variant=variant)
    return work, outs, tape


@pytest.mark.parametrize("size", [(151, 133), (320, 300)])
@pytest.mark.parametrize("variant", range(NV))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("mode", ["tm", "te"])
def test_tiled_matches_two_pass(mode, dtype, variant, size):
    k = 77
    big = size == (320, 300)
    ref_w, ref_o, ref_t = _run(mode, dtype, -1, k, size=size)
    w, o, t = _run(mode, dtype, variant, k, size=size, expect_interior=True if big else None)
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


@pytest.mark.parametrize("variant", range(NV))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_tiled_tm_tfsf_matches_two_pass(dtype, variant):
    """TFSF magnetic currents (the h_masks / h_amplitude protocol of
    run.py:17-23: Hy and Hx sheets, stepper_tm.py:186-191) in the tiled TM
    kernel, on a 320 x 300 grid with interior tiles and sheets crossing tile
    and halo boundaries: fields, psi and spectra bit-identical to the
    two-pass path."""
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_tm import FIELDS, init_tm_state
    nx, ny, k = 320, 300, 53
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(g, eps_inf=_eps(nx, ny, dtype), npml=10)
    prof = GaussianPulse(0.8, 0.5, cutoff=2.0)

    class Tfsf:
        def __init__(self):
            self.mask = torch.zeros(nx, ny, dtype=dtype, device="cuda")
            self.mask[70, 20:250] = 20.0
            hy = torch.zeros(nx - 1, ny, dtype=dtype, device="cuda")
            hy[69, 20:250] = -13.0
            hx = torch.zeros(nx, ny - 1, dtype=dtype, device="cuda")
            hx[30:290, 131] = 7.0
            self.h_masks = {"src_hy_mask": hy, "src_hx_mask": hx}

        def amplitude(self, t):
            return prof.amplitude((t + 0.5) * m.dt)

        def h_amplitude(self, t):
            return -prof.amplitude(t * m.dt - 0.025)

    src = Tfsf()
    n_in, _ = engine2d.tile_counts("tm", nx, ny, m, dtype, variant)
    assert n_in > 0
    out = []
    for v in (-1, variant):
        st = init_tm_state(g)
        work = {f: getattr(st, f).clone() for f in FIELDS}
        plan = engine2d.source_plan(src, (nx, ny), "cuda", dtype, hshapes=engine2d.h_shapes("tm", nx, ny))
        mons = [FluxMonitorTM(g, [0.6, 0.8], x_normal=True, idx=200, lo=15, hi=280)]
        o = engine2d.run_chunk("tm", work, m, 0, k, plan, mons, variant=v)
        out.append((work, o))
    (rw, ro), (w, o) = out
    for f in rw:
        assert torch.equal(w[f], rw[f]), (f, float((w[f] - rw[f]).abs().max()))
    assert torch.equal(o[0], ro[0])
    assert float(rw["Hy"].abs().max()) > 0
