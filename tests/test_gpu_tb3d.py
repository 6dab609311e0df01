"""Two-step 3D pass (variants 4 and 5: two Yee steps per pass over the slab
inside the x and y CPML -- every z column, z walls and z CPML included --
with tile rows 10 or 8 held in registers, plus the two steps of the x and
y CPML cells on a third field set) against the two-pass kernels (variant
0): bit-identical fields, CPML memory and monitor spectra for ragged grids
spanning several tiles, k chunks (psi-free and z-edge ones) and x
segments, point / mask / plane sources (inside the slab and across its
edge), x-normal flux and DFT-field planes (their first-step fields come
from the third set), odd and even step counts; runs the variants do not
take (y-normal monitors, FP64, the E tape) fall back to the two-pass step."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

CASES = [
    # nx, ny, nz, npml, source kind, steps
    (60, 40, 140, 6, "point", 21),
    (45, 70, 75, 5, "mask", 20),
    (90, 33, 66, 6, "plane", 19),
    (70, 64, 130, 8, "point", 16),
]


def _setup(nx, ny, nz, npml, kind, dtype, y_monitor=False):
    from meep_torch_amd import synth
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, FluxMonitor3D
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    r = torch.tensor(synth.splitmix_uniform(nx * ny * nz, nx + 2 * ny + nz).reshape(nx, ny, nz))
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
    elif kind == "mask":   # crosses the interior box's edge in y and z
        m = torch.zeros(nx, ny - 1, nz, dtype=dtype)
        m[nx // 3, 3:ny - 4, 4:nz - 5] = 1.0 / 0.05
        src = Src(("Ey", m.cuda()))
    else:
        plane = torch.zeros(ny, nz - 1, dtype=dtype)
        plane[2:ny - 3, 3:nz - 4] = 1.0
        src = Src(("Ez", nx // 4, plane.cuda()))
    mons = [FluxMonitor3D(c.dt, [0.5, 0.6, 0.7], normal="x", idx=nx - npml - 3, ap=npml + 1, dtype=dtype),
            FluxMonitor3D(c.dt, [0.55, 0.65], normal="x", idx=nx // 2 + 1, ap=2, dtype=dtype),
            DFTFieldsMonitor3D(c.dt, [0.6], "Hy", plane=(0, npml + 4), dtype=dtype)]
    if y_monitor:
        mons.append(FluxMonitor3D(c.dt, [0.55], normal="y", idx=ny // 2, ap=2, dtype=dtype))
    st0 = init_3d_state(nx, ny, nz, npml=npml, dtype=dtype, device="cuda")
    return c, src, mons, st0


def _run(case, dtype, variant, y_monitor=False, tape=False):
    from meep_torch_amd.core.engine3d import run_fields
    from meep_torch_amd.core.stepper_3d import COMPS, PSI_NAMES
    nx, ny, nz, npml, kind, steps = case
    c, src, mons, st0 = _setup(nx, ny, nz, npml, kind, dtype, y_monitor)
    work = {f: getattr(st0, f).clone().contiguous() for f in COMPS + PSI_NAMES}
    tp = None
    if tape:
        n_e = sum(work[f].numel() for f in ("Ex", "Ey", "Ez"))
        tp = torch.zeros(steps * n_e, dtype=dtype, device="cuda")
    outs = run_fields(work, c, steps, src, mons, variant=variant, tape=tp)
    return work, outs, tp


def _same(ref, got, tag):
    (rf, ro, rt), (gf, go, gt) = ref, got
    for f in rf:
        assert torch.equal(gf[f], rf[f]), (tag, f, float((gf[f] - rf[f]).abs().max()))
    assert len(go) == len(ro)
    for a, b in zip(go, ro):
        assert torch.equal(a, b), (tag, "monitor", float((a - b).abs().max()))
    if rt is not None:
        assert torch.equal(gt, rt), (tag, "tape")
    assert float(rf["Ez"].abs().max()) > 0


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}-{c[4]}-{c[5]}" for c in CASES])
def test_two_step_bitexact_vs_two_pass(case):
    ref = _run(case, torch.float32, 0)
    for variant in (4, 5):
        _same(ref, _run(case, torch.float32, variant), variant)


def test_two_step_falls_back():
    case = CASES[0]
    _same(_run(case, torch.float32, 0, y_monitor=True), _run(case, torch.float32, 4, y_monitor=True), "y-normal")
    _same(_run(case, torch.float64, 0), _run(case, torch.float64, 4), "f64")
    _same(_run(case, torch.float32, 0, tape=True), _run(case, torch.float32, 4, tape=True), "tape")
