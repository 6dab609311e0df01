"""Split 3D step (variants 1-3: a one-pass interior kernel streaming along x
with LDS tiles of 6, 14 or 4 rows, plus the two-pass cell updates over the
CPML shell pieces, on ping-pong buffers) against the two-pass kernels
(variant 0): bit-identical fields, CPML memory, monitor spectra and E tape,
for ragged grids spanning several tiles (partial ones, halo rows and
columns) and x segments, point / mask / plane sources, the functional and
lean forms, FP32 and FP64, and odd/even step counts."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

CASES = [
    # nx, ny, nz, npml, source kind, steps
    (30, 24, 20, 6, "point", 41),
    (37, 29, 70, 5, "mask", 30),
    (52, 70, 33, 6, "plane", 27),
    (96, 20, 130, 8, "point", 24),
]


def _setup(nx, ny, nz, npml, kind, dtype):
    from meep_torch_amd import synth
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, FluxMonitor3D
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    r = torch.tensor(synth.splitmix_uniform(nx * ny * nz, nx + ny + nz).reshape(nx, ny, nz))
    eps = (1.0 + 3.0 * r).to(dtype).cuda()
    c = build_3d_coeffs(nx, ny, nz, 0.05, eps=eps, npml=npml, dtype=dtype, device="cuda")
    prof = GaussianPulse(0.6, 0.3)

    class Src:
        def __init__(self, tup):
            self.src = tup

        def amplitude(self, t):
            return prof.amplitude((t + 0.5) * c.dt)

    if kind == "point":
        src = Src(("Ez", nx // 2, ny // 2, nz // 3))
    elif kind == "mask":
        m = torch.zeros(nx, ny - 1, nz, dtype=dtype)
        m[nx // 3, 3:ny - 4, 4:nz - 5] = 1.0 / 0.05
        src = Src(("Ey", m.cuda()))
    else:
        plane = torch.zeros(ny, nz - 1, dtype=dtype)
        plane[2:ny - 3, 3:nz - 4] = 1.0
        src = Src(("Ez", nx // 4, plane.cuda()))
    mons = [FluxMonitor3D(c.dt, [0.5, 0.6, 0.7], normal="x", idx=nx - npml - 3, ap=npml + 1, dtype=dtype),
            FluxMonitor3D(c.dt, [0.55, 0.65], normal="y", idx=ny // 2, ap=2, dtype=dtype),
            DFTFieldsMonitor3D(c.dt, [0.6], "Hy", plane=(2, nz // 2), dtype=dtype)]
    st0 = init_3d_state(nx, ny, nz, npml=npml, dtype=dtype, device="cuda")
    return c, src, mons, st0


def _run(case, dtype, variant, lean):
    from meep_torch_amd.core.engine3d import run_fields
    from meep_torch_amd.core.stepper_3d import COMPS, PSI_NAMES
    nx, ny, nz, npml, kind, steps = case
    c, src, mons, st0 = _setup(nx, ny, nz, npml, kind, dtype)
    work = {f: getattr(st0, f).clone().contiguous() for f in COMPS + PSI_NAMES}
    outs = run_fields(work, c, steps, src, mons, lean=lean, variant=variant)
    return work, outs


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}-{c[4]}" for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_fused_bitexact_vs_two_pass(case, dtype):
    for lean in (False, True):
        ref, ref_out = _run(case, dtype, 0, lean)
        for variant in (1, 2, 3):
            got, out = _run(case, dtype, variant, lean)
            for f in ref:
                assert torch.equal(got[f], ref[f]), (variant, lean, f, float((got[f] - ref[f]).abs().max()))
            for a, b in zip(out, ref_out):
                assert torch.equal(a, b), (variant, lean, "monitor")
            assert float(ref["Ez"].abs().max()) > 0


def test_fused_tape_matches_two_pass():
    from meep_torch_amd.core.engine3d import run_fields
    from meep_torch_amd.core.stepper_3d import COMPS, PSI_NAMES
    case = CASES[1]
    nx, ny, nz, npml, kind, steps = case
    tapes = []
    for variant in (0, 1, 2, 3):
        c, src, mons, st0 = _setup(nx, ny, nz, npml, kind, torch.float64)
        work = {f: getattr(st0, f).clone().contiguous() for f in COMPS + PSI_NAMES}
        slot = sum(work[f].numel() for f in ("Ex", "Ey", "Ez"))
        tape = torch.zeros(steps * slot, dtype=torch.float64, device="cuda")
        run_fields(work, c, steps, src, [], variant=variant, tape=tape)
        tapes.append(tape)
    for t in tapes[1:]:
        assert torch.equal(tapes[0], t)
