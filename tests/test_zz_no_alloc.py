"""The allocation-free TM entry point (mt_tm2d_run_static: DFT staging in
the library's static device array, no workspace argument) against
mt_tm2d_run with a caller workspace: fields, CPML memory and spectra bit
for bit, on two calls in a row -- the second on a smaller grid with fewer
monitor elements, over what the first left in the shared array -- for the
two-pass path and a tiled variant; and the size limit's error."""
import ctypes

import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def _case(nx, ny, variant, static):
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import DFTFieldsMonitor, FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import FIELDS, init_tm_state
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=torch.float32, device="cuda")
    x = torch.arange(nx, dtype=torch.float64)[:, None]
    eps = (1.0 + 2.0 * torch.exp(-((x - nx / 2) ** 2) / 300.0)).expand(nx, ny).contiguous()
    m = build_tm_coeffs(g, eps_inf=eps.float().cuda(), npml=10)
    src = LineSource(g, ix=15, j0=12, j1=ny - 12, profile=GaussianPulse(0.8, 0.5, cutoff=2.0))
    mons = [FluxMonitorTM(g, [0.6, 0.8, 1.0], x_normal=True, idx=50, lo=11, hi=ny - 11),
            DFTFieldsMonitor(g, [0.9], component="Ez")]
    st = init_tm_state(g)
    work = {f: getattr(st, f).clone() for f in FIELDS}
    plan = engine2d.source_plan(src, (nx, ny), "cuda", torch.float32)
    outs = engine2d.run_chunk("tm", work, m, 0, 300, plan, mons, variant=variant, static_stage=static)
    torch.cuda.synchronize()
    return work, outs


@pytest.mark.parametrize("variant", [-1, 5])
def test_static_staging_equals_workspace_staging(variant):
    for nx, ny in ((260, 230), (150, 120)):
        ref_w, ref_o = _case(nx, ny, variant, False)
        w, o = _case(nx, ny, variant, True)
        for f in ref_w:
            assert torch.equal(w[f], ref_w[f]), (nx, f)
        for a, b in zip(o, ref_o):
            assert torch.equal(a, b), nx
        assert float(ref_o[0].abs().max()) > 0


def test_static_staging_rejects_oversized_monitors():
    """A DFT-field monitor over a grid too large for the static array raises
    a clear error instead of allocating."""
    from meep_torch_amd import _engine as E
    from meep_torch_amd.core import engine2d
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import DFTFieldsMonitor
    from meep_torch_amd.core.stepper_tm import FIELDS, init_tm_state
    cap = E.lib().mt_static_stage_bytes()
    n = int((cap / 4) ** 0.5) + 64                # one FP32 step of Ez over n^2 cells exceeds the array
    g = Grid(nx=n, ny=n, dx=0.05, dtype=torch.float32, device="cuda")
    m = build_tm_coeffs(g, eps_inf=1.0, npml=4)
    st = init_tm_state(g)
    work = {f: getattr(st, f).clone() for f in FIELDS}
    with pytest.raises(RuntimeError, match="static staging array"):
        engine2d.run_chunk("tm", work, m, 0, 2, None, [DFTFieldsMonitor(g, [0.5], component="Ez")],
                           variant=-1, static_stage=True)
