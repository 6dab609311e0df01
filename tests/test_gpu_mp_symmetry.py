"""Symmetries in the 2D Simulation facade (meep_torch/mp/simulation.py:
383-494 fold rules, :822-837 source folding, :874-970 full-cell
reconstruction, :1829-1843 flux folding) against the reference's own
Simulation runs (golden sim_sym_f64): Mirror(Y) on a guide with a decay stop
probed below the plane, Mirror(X, -1) + Mirror(Y) with a source pair,
Rotate2, Rotate4 and a TE Mirror(Y, -1).  The run keeps the part of the cell
at and above the cut plane(s) on the folded core stepper; the full-cell
field, the fluxes (lines crossing a plane double back, lines below it flip
sign), the probed points (folded with the Ez phase) and the stop time equal
the reference's within FP64 rounding."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def _cases(mp):
    guide = [mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))]
    pulse = lambda f: mp.GaussianSource(f, fwidth=0.6)
    return {
        "my": (dict(cell_size=mp.Vector3(6, 4), geometry=guide,
                    sources=[mp.Source(pulse(0.8), component=mp.Ez, center=mp.Vector3(-2.0, 0.0),
                                       size=mp.Vector3(0, 0.8))],
                    symmetries=[mp.Mirror(mp.Y)]),
               [mp.FluxRegion(center=mp.Vector3(2.0, 0), size=mp.Vector3(0, 3)),
                mp.FluxRegion(center=mp.Vector3(0.5, -1.0), size=mp.Vector3(2.0, 0))],
               [(1.0, 0.3), (1.0, -0.3)],
               dict(until_after_sources=mp.stop_when_fields_decayed(2.0, mp.Ez, mp.Vector3(1.0, -0.3), 1e-2))),
        "mxy": (dict(cell_size=mp.Vector3(5, 4), geometry=[],
                     sources=[mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(0.8, 0.0)),
                              mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(-0.8, 0.0), amplitude=-1.0)],
                     symmetries=[mp.Mirror(mp.X, phase=-1), mp.Mirror(mp.Y)]),
                [mp.FluxRegion(center=mp.Vector3(0, 1.2), size=mp.Vector3(4.0, 0)),
                 mp.FluxRegion(center=mp.Vector3(1.8, 0), size=mp.Vector3(0, 3.0))],
                [(0.5, 0.5), (-0.5, -0.5), (-1.2, 0.7)], dict(until=8)),
        "r2": (dict(cell_size=mp.Vector3(4, 4), geometry=[mp.Block(size=mp.Vector3(1.2, 0.4),
                                                                   material=mp.Medium(epsilon=6.0))],
                    sources=[mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(0.5, 0.5)),
                             mp.Source(pulse(0.9), component=mp.Ez, center=mp.Vector3(-0.5, -0.5))],
                    symmetries=[mp.Rotate2(mp.Z)]),
               [mp.FluxRegion(center=mp.Vector3(1.2, 0.9), size=mp.Vector3(0, 1.4))],
               [(0.6, 0.4), (-0.6, -0.4), (0.9, -0.8)], dict(until=8)),
        "r4": (dict(cell_size=mp.Vector3(4, 4), geometry=[mp.Block(size=mp.Vector3(0.8, 0.8),
                                                                   material=mp.Medium(epsilon=4.0))],
                    sources=[mp.Source(pulse(1.0), component=mp.Ez, center=mp.Vector3(0.6, 0.7))],
                    symmetries=[mp.Rotate4(mp.Z)]),
               [mp.FluxRegion(center=mp.Vector3(1.4, 0.9), size=mp.Vector3(0, 1.4))],
               [(0.6, 0.7), (-0.7, 0.6), (-0.6, -0.7)], dict(until=8)),
        "te": (dict(cell_size=mp.Vector3(6, 4), geometry=guide,
                    sources=[mp.Source(pulse(0.8), component=mp.Ey, center=mp.Vector3(-2.0, 0.0),
                                       size=mp.Vector3(0, 0.8))],
                    symmetries=[mp.Mirror(mp.Y, phase=-1)]),
               [mp.FluxRegion(center=mp.Vector3(2.0, 0), size=mp.Vector3(0, 3))],
               [(1.0, 0.3)], dict(until=10)),
    }


@pytest.mark.parametrize("name", ["my", "mxy", "r2", "r4", "te"])
def test_symmetric_runs_match_reference(name):
    import meep_torch_amd.mp as mp
    G = golden("sim_sym_f64")
    kw, regions, pts, runkw = _cases(mp)[name]
    print(f"[sym] {name}", flush=True)
    sim = mp.Simulation(resolution=10, boundary_layers=[mp.PML(0.5)], dtype=torch.float64, device="cuda", **kw)
    fls = [sim.add_flux(0.8, 0.4, 3, r) for r in regions]
    comp = mp.Ez if name != "te" else mp.Hz
    seen = []
    sim.run(mp.at_every(1.0, lambda s: seen.append([float(s.get_field_point(comp, mp.Vector3(*p))) for p in pts])),
            **runkw)
    assert sim.grid.nx * sim.grid.ny < sim.gs.nx * sim.gs.ny          # the run used the folded domain
    assert abs(sim.meep_time() - float(G[f"{name}_t"])) < 1e-9
    f, ref = np.asarray(sim.get_array(component=comp)), G[f"{name}_field"]
    assert f.shape == ref.shape
    assert np.linalg.norm(f - ref) <= 1e-12 * np.linalg.norm(ref)
    fx = np.array([mp.get_fluxes(fl) for fl in fls])
    assert np.abs(fx - G[f"{name}_flux"]).max() <= 1e-10 * np.abs(G[f"{name}_flux"]).max()
    assert np.allclose(np.array(seen), G[f"{name}_probe"], rtol=1e-10, atol=1e-14 * np.abs(G[f"{name}_probe"]).max())
