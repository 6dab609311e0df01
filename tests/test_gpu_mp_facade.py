"""Simulation facade end to end on the HIP engine: the GOLDEN_PLAIN scene
(100x60 TM guide, eps 12, 10-cell PML, 800 steps, 3 flux frequencies) against
the reference's own run, in FP32 (fields <= 1e-5 rel-L2, spectra <= 1e-4) and
FP64 (<= 1e-10); consecutive runs integrate to the accumulated clock like one
continuing run; DFT field arrays match a run_tm DFT monitor."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def _plain(dtype):
    import meep_torch_amd.mp as mp
    sim = mp.Simulation(
        cell_size=mp.Vector3(10, 6), resolution=10,
        geometry=[mp.Block(center=mp.Vector3(), size=mp.Vector3(mp.inf, 1, mp.inf),
                           material=mp.Medium(epsilon=12))],
        sources=[mp.Source(mp.GaussianSource(0.15, fwidth=0.1), component=mp.Ez,
                           center=mp.Vector3(-3, 0), size=mp.Vector3(0, 2))],
        boundary_layers=[mp.PML(1.0)], dtype=dtype, device="cuda")
    fl = sim.add_flux(0.15, 0.1, 3, mp.FluxRegion(center=mp.Vector3(3, 0), size=mp.Vector3(0, 4)))
    return mp, sim, fl


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dtype,tag,tol_f,tol_s", [(torch.float32, "f32", 1e-5, 1e-4),
                                                    (torch.float64, "f64", 1e-10, 1e-10)])
def test_golden_plain(dtype, tag, tol_f, tol_s):
    g = golden("golden_plain")
    mp, sim, fl = _plain(dtype)
    sim.run(until=40)
    flux = np.array(mp.get_fluxes(fl))
    ez = sim.get_array(component=mp.Ez)
    e_flux, e_ez = _rel(flux, g[f"{tag}_flux"]), _rel(ez, g[f"{tag}_Ez"])
    print(f"\n[golden_plain {tag}] flux rel {e_flux:.2e}, Ez rel {e_ez:.2e}; "
          f"vs f64 reference: flux {_rel(flux, g['f64_flux']):.2e}, Ez {_rel(ez, g['f64_Ez']):.2e}")
    assert e_flux <= tol_s and e_ez <= tol_f
    assert np.array_equal(sim.get_array(component=mp.Dielectric), g[f"{tag}_eps"])


def test_consecutive_runs_accumulate_clock():
    mp, a, fa = _plain(torch.float64)
    a.run(until=20)
    a.run(until=20)
    mp, b, fb = _plain(torch.float64)
    b.run(until=40)
    assert a.round_time() == b.round_time() == 40.0
    assert np.allclose(a.get_array(component=mp.Ez), b.get_array(component=mp.Ez), rtol=0, atol=1e-13)
    assert np.allclose(mp.get_fluxes(fa), mp.get_fluxes(fb), rtol=1e-12, atol=0)


def test_dft_fields_match_run_tm_monitor():
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.monitors import DFTFieldsMonitor
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    _, sim, _ = _plain(torch.float64)
    d = sim.add_dft_fields([mp.Ez], 0.15, 0.1, 2, center=mp.Vector3(1, 0), size=mp.Vector3(2, 2))
    sim.run(until=10)
    arr = sim.get_dft_array(d, mp.Ez, 1)
    mon = DFTFieldsMonitor(sim.grid, d.freqs, component="Ez")
    src = sim._make_sources()[0]
    _, (acc,) = run_tm(init_tm_state(sim.grid), sim._coeffs, 200, source=src, monitors=[mon])
    ref = (acc[1, 0] + 1j * acc[1, 1]).cpu().numpy()
    i0, j0 = 50, 20
    assert arr.shape == (20, 20)
    assert np.array_equal(arr, ref[i0:i0 + 20, j0:j0 + 20])
