"""3D Simulation facade on the HIP engine against a hand-built run_3d call
(meep_torch/mp/simulation.py:630-766 is an adapter over run_3d): a 32 x 24 x 20
cell at resolution 10 with a 4-cell PML, one eps-4 Block voxel-filled over an
index box worked out by hand, a point Ez dipole of amplitude 2 and a
full-cell x-normal flux plane (aperture npml + 1).  The facade's eps array,
fields and fluxes must equal the direct call exactly; a second run restarts
the fields (same result) while the clock keeps accumulating, as in the
reference."""
import numpy as np
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


class _Scaled:
    def __init__(self, src, a):
        self.src, self._s, self._a = src.src, src, a

    def amplitude(self, t):
        return self._a * self._s.amplitude(t)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_3d_facade_equals_direct_run(dtype):
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    gsrc = mp.GaussianSource(0.8, fwidth=0.4)
    sim = mp.Simulation(cell_size=mp.Vector3(3.2, 2.4, 2.0), resolution=10, dimensions=3,
                        geometry=[mp.Block(center=mp.Vector3(0, 0, -0.5), size=mp.Vector3(mp.inf, 0.6, 0.4),
                                           material=mp.Medium(epsilon=4))],
                        sources=[mp.Source(gsrc, component=mp.Ez, center=mp.Vector3(-1.0, 0, -0.5), amplitude=2.0)],
                        boundary_layers=[mp.PML(0.4)], dtype=dtype, device="cuda")
    fl = sim.add_flux(0.8, 0.4, 5, mp.FluxRegion(center=mp.Vector3(1.0, 0, 0), size=mp.Vector3(0, mp.inf, mp.inf)))
    sim.run(until=5)

    # by hand: 32 x 24 x 20 cells; the block covers y in [0.9, 1.5] um -> j 9..15,
    # z in [0.3, 0.7] um -> k 3..7, all x; the dipole sits at (6, 12, 5); the
    # plane at x = 2.6 um -> i = 26; dt = 0.5 * 0.1 / sqrt(3); until 5 -> 173 steps
    eps = torch.ones(32, 24, 20, dtype=dtype, device="cuda")
    eps[:, 9:15, 3:7] = 4.0
    c = build_3d_coeffs(32, 24, 20, 0.1, eps=eps, npml=4, dtype=dtype, device="cuda")
    n_steps = int(round(5 / c.dt))
    assert n_steps == 173
    src = _Scaled(PointSource3D(c.dt, "Ez", 6, 12, 5, gsrc.profile()), 2.0)
    freqs = np.linspace(0.8 - 0.4 / 2.0, 0.8 + 0.4 / 2.0, 5).tolist()   # meep's fcen +- df/2 grid
    mon = FluxMonitor3D(c.dt, freqs, normal="x", idx=26, ap=5, dtype=dtype)
    st, (acc,) = run_3d(init_3d_state(32, 24, 20, npml=4, dtype=dtype, device="cuda"), c, n_steps,
                        source=src, monitors=[mon])
    ref_flux = mon.flux(acc, 0.1).detach().cpu().numpy()

    assert np.array_equal(sim.get_array(component=mp.Dielectric), eps.cpu().numpy())
    assert np.array_equal(sim.get_array(component=mp.Ez), st.Ez.cpu().numpy())
    assert np.array_equal(np.asarray(mp.get_fluxes(fl)), ref_flux)
    # the monitor keeps its frequencies in the run's dtype
    assert np.allclose(mp.get_flux_freqs(fl), freqs, rtol=0, atol=1e-15 if dtype == torch.float64 else 1e-7)
    assert float(np.abs(ref_flux).max()) > 0
    assert abs(sim.round_time() - 173 * c.dt) < 1e-12

    # the reference restarts 3D fields on every run and keeps counting time
    sim.run(until=5)
    assert np.array_equal(sim.get_array(component=mp.Ez), st.Ez.cpu().numpy())
    assert abs(sim.round_time() - 2 * 173 * c.dt) < 1e-12


def test_3d_facade_plane_source_and_dft_fields():
    """A windowed plane source (finite transverse size -> aperture [npml,
    min(n) - npml)) and a full-cell DFT field monitor, against the direct call."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, PlaneSource, run_3d
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    gsrc = mp.GaussianSource(0.8, fwidth=0.4)
    sim = mp.Simulation(cell_size=mp.Vector3(2.4, 2.0, 2.0), resolution=10, dimensions=3,
                        sources=[mp.Source(gsrc, component=mp.Ey, center=mp.Vector3(-0.6, 0, 0),
                                           size=mp.Vector3(0, 1.0, 1.0))],
                        boundary_layers=[mp.PML(0.3)], dtype=torch.float64, device="cuda")
    dft = sim.add_dft_fields([mp.Ey], 0.8, 0.0, 1)
    sim.run(until=3)
    c = build_3d_coeffs(24, 20, 20, 0.1, eps=torch.ones(24, 20, 20, dtype=torch.float64, device="cuda"), npml=3,
                        dtype=torch.float64, device="cuda")
    # plane normal to x (the zero extent) at i = (−0.6 + 1.2) / 0.1 = 6, window [3, 17)
    src = PlaneSource(24, 20, 20, c.dt, "Ey", 0, 6, gsrc.profile(), lo=3, hi=17, device="cuda")

    class _One:
        def __init__(self, s):
            self.src, self._s = s.src, s

        def amplitude(self, t):
            return 1.0 * self._s.amplitude(t)

    mon = DFTFieldsMonitor3D(c.dt, [0.8], component="Ey", dtype=torch.float64)
    _, (acc,) = run_3d(init_3d_state(24, 20, 20, npml=3, dtype=torch.float64, device="cuda"), c,
                       int(round(3 / c.dt)), source=_One(src), monitors=[mon])
    a = acc[0].cpu().numpy()
    got = sim.get_dft_array(dft, mp.Ey, 0)
    assert np.array_equal(got, a[0] + 1j * a[1])
    assert float(np.abs(got).max()) > 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_3d_facade_media_equal_direct_run(dtype):
    """A Block of conductive Drude metal (D_conductivity plus one Drude pole)
    and a slab with B_conductivity: the facade's run equals run_3d on
    sigma_d / sigma_b maps and a pole mask built from the painted cells."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    gsrc = mp.GaussianSource(0.8, fwidth=0.4)
    drude = mp.DrudeSusceptibility(frequency=1.2, gamma=0.2, sigma=2.0)
    metal = mp.Medium(epsilon=2, D_conductivity=0.5, E_susceptibilities=[drude])
    lossy = mp.Medium(epsilon=3, B_conductivity=0.3)
    sim = mp.Simulation(cell_size=mp.Vector3(3.2, 2.4, 2.0), resolution=10, dimensions=3,
                        geometry=[mp.Block(center=mp.Vector3(0.2, 0, 0), size=mp.Vector3(1.0, 0.8, 0.6),
                                           material=metal),
                                  mp.Block(center=mp.Vector3(0.6, 0, 0), size=mp.Vector3(0.2, 1.2, 1.0),
                                           material=lossy)],
                        sources=[mp.Source(gsrc, component=mp.Ez, center=mp.Vector3(-1.0, 0, 0), amplitude=2.0)],
                        boundary_layers=[mp.PML(0.4)], dtype=dtype, device="cuda")
    fl = sim.add_flux(0.8, 0.4, 5, mp.FluxRegion(center=mp.Vector3(1.0, 0, 0), size=mp.Vector3(0, mp.inf, mp.inf)))
    sim.run(until=5)

    eps = torch.from_numpy(sim.get_array(component=mp.Dielectric)).to(device="cuda", dtype=dtype)
    block, slab = (eps == 2.0).to(dtype), (eps == 3.0).to(dtype)
    assert float(block.sum()) > 0 and float(slab.sum()) > 0
    c = build_3d_coeffs(32, 24, 20, 0.1, eps=eps, npml=4, sigma_d=0.5 * block, sigma_b=0.3 * slab,
                        poles=[drude.to_pole()], pole_masks=[block], dtype=dtype, device="cuda")
    n_steps = int(round(5 / c.dt))
    src = _Scaled(PointSource3D(c.dt, "Ez", 6, 12, 10, gsrc.profile()), 2.0)
    freqs = np.linspace(0.8 - 0.4 / 2.0, 0.8 + 0.4 / 2.0, 5).tolist()
    mon = FluxMonitor3D(c.dt, freqs, normal="x", idx=26, ap=5, dtype=dtype)
    st, (acc,) = run_3d(init_3d_state(32, 24, 20, npml=4, n_poles=1, dtype=dtype, device="cuda"), c, n_steps,
                        source=src, monitors=[mon])
    assert np.array_equal(sim.get_array(component=mp.Ez), st.Ez.cpu().numpy())
    assert np.array_equal(np.asarray(mp.get_fluxes(fl)), mon.flux(acc, 0.1).detach().cpu().numpy())
    assert float(st.Px_cur.abs().max()) > 0


@pytest.mark.parametrize("kind", ["point", "plane"])
def test_3d_facade_matches_reference_simulation(kind):
    """The facade's 3D source mapping against the reference's own 3D
    Simulation (golden sim3d_f64, written by the reference's mp.Simulation,
    mp/simulation.py:630-766): a point Ez dipole, and a finite-aperture Ey
    plane source (whose sign and aperture differ from the point form,
    stepper_3d.py:452-481), each with a voxel Block and a flux plane; E
    fields and fluxes within FP64 rounding."""
    import meep_torch_amd.mp as mp
    from conftest import golden
    G = golden("sim3d_f64")
    cell = mp.Vector3(1.6, 1.4, 1.2)
    geom = [mp.Block(center=mp.Vector3(0.1, 0, -0.1), size=mp.Vector3(mp.inf, 0.3, 0.25),
                     material=mp.Medium(epsilon=6.0))]
    if kind == "point":
        srcs = [mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ez, center=mp.Vector3(-0.4, 0.05, 0.1))]
    else:
        srcs = [mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ey, center=mp.Vector3(-0.4, 0, 0),
                          size=mp.Vector3(0, 0.6, 0.5))]
    sim = mp.Simulation(cell_size=cell, resolution=20, geometry=geom, sources=srcs,
                        boundary_layers=[mp.PML(0.2)], dimensions=3, dtype=torch.float64, device="cuda")
    fl = sim.add_flux(1.0, 0.4, 3, mp.FluxRegion(center=mp.Vector3(0.4, 0, 0), size=mp.Vector3(0, 1.4, 1.2)))
    sim.run(until=6.0)
    for name, comp in (("Ex", mp.Ex), ("Ey", mp.Ey), ("Ez", mp.Ez)):
        ref = G[f"{kind}_{name}"]
        got = np.asarray(sim.get_array(component=comp))
        assert got.shape == ref.shape, (name, got.shape, ref.shape)
        assert np.linalg.norm(got - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300), name
    assert float(np.abs(G[f"{kind}_Ey" if kind == "plane" else f"{kind}_Ez"]).max()) > 0
    ref_flux = G[f"{kind}_flux"]
    got_flux = np.asarray(mp.get_fluxes(fl))
    assert np.abs(got_flux - ref_flux).max() <= 1e-10 * np.abs(ref_flux).max(), (got_flux, ref_flux)


def test_3d_facade_absorber_matches_reference_simulation():
    """mp.Absorber all round a 3D cell: no CPML, the matched-sigma absorber
    path (build_3d_coeffs(npml=0, n_abs=...), stepper_3d.py:328-361) with a
    plane source and a flux plane; Ey and the flux equal the reference
    Simulation's (golden sim_absorber_f64) within FP64 rounding."""
    import meep_torch_amd.mp as mp
    from conftest import golden
    G = golden("sim_absorber_f64")
    sim = mp.Simulation(cell_size=mp.Vector3(1.6, 1.4, 1.2), resolution=20,
                        geometry=[mp.Block(center=mp.Vector3(0.1, 0, -0.1), size=mp.Vector3(mp.inf, 0.3, 0.25),
                                           material=mp.Medium(epsilon=6.0))],
                        sources=[mp.Source(mp.GaussianSource(1.0, fwidth=1.0), component=mp.Ey,
                                           center=mp.Vector3(-0.4, 0, 0), size=mp.Vector3(0, 0.6, 0.5))],
                        boundary_layers=[mp.Absorber(0.25)], dimensions=3, dtype=torch.float64, device="cuda")
    fl = sim.add_flux(1.0, 0.4, 3, mp.FluxRegion(center=mp.Vector3(0.4, 0, 0), size=mp.Vector3(0, 1.0, 0.8)))
    sim.run(until=6.0)
    assert sim._coeffs.cpml is None
    got, ref = np.asarray(sim.get_array(component=mp.Ey)), G["d3_Ey"]
    assert got.shape == ref.shape
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    fx = np.asarray(mp.get_fluxes(fl))
    assert np.abs(fx - G["d3_flux"]).max() <= 1e-10 * np.abs(G["d3_flux"]).max()
