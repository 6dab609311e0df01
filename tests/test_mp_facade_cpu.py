"""Host side of the meep-compatible Simulation facade (no GPU work): the
GOLDEN_PLAIN scene (meep_torch tests/mp_parity/test_meep_flow.py:49-54)
rasterizes to the reference's eps array bit-exactly, its source, monitor and
PML indices follow meep_torch/mp/coords.py, flux frequencies equal the
reference's (in each dtype), Block coverage is the exact area fraction, and
unsupported options raise NotImplementedError."""
import numpy as np
import pytest
import torch

from conftest import golden


def _plain(dtype):
    import meep_torch_amd.mp as mp
    return mp, mp.Simulation(
        cell_size=mp.Vector3(10, 6), resolution=10,
        geometry=[mp.Block(center=mp.Vector3(), size=mp.Vector3(mp.inf, 1, mp.inf),
                           material=mp.Medium(epsilon=12))],
        sources=[mp.Source(mp.GaussianSource(0.15, fwidth=0.1), component=mp.Ez,
                           center=mp.Vector3(-3, 0), size=mp.Vector3(0, 2))],
        boundary_layers=[mp.PML(1.0)], dtype=dtype, device="cpu")


@pytest.mark.parametrize("dtype,tag", [(torch.float32, "f32"), (torch.float64, "f64")])
def test_golden_plain_setup_matches_reference(dtype, tag):
    g = golden("golden_plain")
    mp, sim = _plain(dtype)
    eps = sim.get_array(component=mp.Dielectric)
    assert eps.dtype == np.float64 and np.array_equal(eps, g[f"{tag}_eps"])
    fl = sim.add_flux(0.15, 0.1, 3, mp.FluxRegion(center=mp.Vector3(3, 0), size=mp.Vector3(0, 4)))
    assert np.array_equal(np.array(mp.get_flux_freqs(fl)), g[f"{tag}_freqs"])
    mon = sim._monitors[fl.name][1]
    assert (mon.idx, mon.lo, mon.hi, mon.x_normal) == (80, 10, 50, True)
    src = sim._make_sources()[0]
    nz = src.mask.nonzero()
    assert nz[:, 0].unique().tolist() == [20] and nz[:, 1].tolist() == list(range(20, 40))
    assert sim._npml_sides == [[10, 10], [10, 10]]
    assert int(round(40 / sim.grid.dt)) == 800


def test_block_coverage_is_exact_area_fraction():
    from meep_torch_amd.mp import geometry as G
    from meep_torch_amd.mp.vec import Vector3
    gs = G.grid_spec(Vector3(2, 1), 10)
    b = G.Block(size=Vector3(0.55, 0.33), center=Vector3(0.123, -0.07), material=G.Medium(epsilon=4))
    cov = b.coverage(gs)
    x0, y0, x1, y1 = G.rect_corners(b.center, b.size, gs)
    assert abs(cov.sum() * gs.dx ** 2 - (x1 - x0) * (y1 - y0)) < 1e-12
    # brute force: 200x200 sub-samples per cell on a few cells
    for (i, j) in [(int(x0 / gs.dx), int(y0 / gs.dx)), (int(x1 / gs.dx), int(y1 / gs.dx)), (10, 3)]:
        s = (np.arange(200) + 0.5) / 200 * gs.dx
        X, Y = np.meshgrid(i * gs.dx + s, j * gs.dx + s, indexing="ij")
        bf = ((X >= x0) & (X < x1) & (Y >= y0) & (Y < y1)).mean()
        assert abs(cov[i, j] - bf) < 0.01


def test_te_kottke_pair_on_interface():
    from meep_torch_amd.mp import geometry as G
    from meep_torch_amd.mp.vec import Vector3
    gs = G.grid_spec(Vector3(2, 2), 10)
    # slab with its edge inside a cell along x: normal is +-x there
    from meep_torch_amd.mp.vec import inf
    b = G.Block(size=Vector3(0.55, inf), center=Vector3(0.0, 0.0),
                material=G.Medium(epsilon=9))
    eps_tm, xx, yy = G.rasterize([b], gs, 1.0)
    i = int(np.argmax((eps_tm[:, 10] > 1) & (eps_tm[:, 10] < 9)))
    c = (eps_tm[i, 10] - 1.0) / 8.0
    harm = 1.0 / (c / 9.0 + (1 - c) / 1.0)
    assert abs(xx[i, 10] - harm) < 1e-12        # E normal to the interface: harmonic mean
    assert abs(yy[i, 10] - eps_tm[i, 10]) < 1e-12  # tangential: arithmetic mean


def test_unsupported_options_raise():
    import meep_torch_amd.mp as mp
    with pytest.raises(NotImplementedError):
        mp.Simulation(cell_size=mp.Vector3(4, 4), resolution=10, symmetries=[object()], device="cpu")
    with pytest.raises(NotImplementedError):
        mp.Simulation(cell_size=mp.Vector3(4, 4, 4), resolution=10, dimensions=3, Courant=0.6, device="cpu")
    with pytest.raises(NotImplementedError):
        mp.Medium(epsilon=2, H_susceptibilities=[object()])
    with pytest.raises(NotImplementedError):   # a cylinder whose axis leaves z has no 2D cross-section here
        mp.Simulation(cell_size=mp.Vector3(4, 4), resolution=10, device="cpu",
                      geometry=[mp.Cylinder(radius=0.5, axis=mp.Vector3(1, 0, 0))]).get_array(component=mp.Dielectric)
    sim3 = mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=3, device="cpu")
    with pytest.raises(NotImplementedError):
        sim3.run(lambda s: None, until=1)        # step functions drive 2D runs only


def test_run_requires_hip_device():
    """No CPU fallback: running the facade on CPU tensors raises (2D and 3D)."""
    _, sim = _plain(torch.float32)
    with pytest.raises(RuntimeError):
        sim.run(until=1)
    import meep_torch_amd.mp as mp
    sim3 = mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=3, device="cpu",
                         sources=[mp.Source(mp.GaussianSource(1.0, fwidth=0.5), component=mp.Ez)])
    with pytest.raises(RuntimeError):
        sim3.run(until=1)


def test_3d_host_setup_closed_forms():
    """3D grid, index boxes and CPML sides by hand (meep_torch/mp/coords.py:34-111,
    mp/simulation.py:89-116,359-379): cell 4 x 3 x 2 at resolution 10."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.mp.sim3d import box3d, cell_index3, grid_spec3, pml_sides3
    gs = grid_spec3(mp.Vector3(4, 3, 2), 10)
    assert (gs.nx, gs.ny, gs.nz) == (40, 30, 20)
    # x in [2.0, 3.0] um from the corner, y the whole axis, z in [0.8, 1.2]
    assert box3d(mp.Vector3(0.5, 0, 0), mp.Vector3(1, mp.inf, 0.4), gs) == (20, 30, 0, 30, 8, 12)
    # clipped at the high x face: [3.6, 4.2] -> [3.6, 4.0]
    assert box3d(mp.Vector3(1.9, 0, 0), mp.Vector3(0.6, 0.2, 0.2), gs)[:2] == (36, 40)
    # a zero-thickness box keeps one cell at the low face
    assert box3d(mp.Vector3(-2, 0, 0), mp.Vector3(0, 0.2, 0.2), gs)[:2] == (0, 1)
    assert cell_index3(mp.Vector3(-2, 0, 0.04), gs) == (0, 15, 10)
    assert pml_sides3([mp.PML(0.5)], gs, mp.Vector3(4, 3, 2)) == [[5, 5], [5, 5], [5, 5]]
    sides = pml_sides3([mp.PML(0.3, direction=mp.X, side=mp.Low), mp.PML(0.2)], gs, mp.Vector3(4, 3, 2))
    assert sides == [[3, 2], [2, 2], [2, 2]]
    assert pml_sides3([mp.PML(0.4, direction=mp.Z, side=mp.High)], gs, mp.Vector3(4, 3, 2)) == \
        [[0, 0], [0, 0], [0, 4]]


def test_3d_courant_and_options_raise():
    import pytest
    import meep_torch_amd.mp as mp
    with pytest.raises(NotImplementedError):
        mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=3, Courant=0.4)
    with pytest.raises(NotImplementedError):
        mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=1)


def test_media_rasterization_compositing():
    """Painter rules of the media maps (meep_torch/geometry/scene.py:84-96,
    meep_torch/mp/geometry.py:265-286): masks are scaled by (1 - coverage)
    under every later object and gain the coverage of objects sharing the
    pole; chi3 blends over objects with nonzero chi3 and a nonzero background
    chi3 is painted last over everything; conductivity / mu blend over every
    object once any medium has one."""
    import numpy as np
    import meep_torch_amd.mp as mp
    from meep_torch_amd.mp import geometry as G
    gs = G.grid_spec(mp.Vector3(4, 4), 10)
    L = mp.LorentzianSusceptibility(frequency=1.2, gamma=0.1, sigma=1.5)
    D = mp.DrudeSusceptibility(frequency=0.9, gamma=0.3, sigma=0.5)
    a = mp.Block(center=mp.Vector3(-0.5, 0), size=mp.Vector3(1.0, mp.inf),
                 material=mp.Medium(epsilon=2, E_susceptibilities=[L]))
    b = mp.Block(center=mp.Vector3(-0.275, 0), size=mp.Vector3(1.0, 2.0),    # half cells on its x edges
                 material=mp.Medium(epsilon=3, E_susceptibilities=[L], chi3=0.1))
    c = mp.Block(center=mp.Vector3(1.0, 0), size=mp.Vector3(0.6, 0.6),
                 material=mp.Medium(E_susceptibilities=[D], D_conductivity=0.3, mu=2.0))
    ca, cb, cc = (o.coverage(gs) for o in (a, b, c))
    assert ((cb > 0) & (cb < 1)).any()                                     # partial cells exist
    out = G.rasterize_media([a, b, c], gs, mp.Medium())
    assert out["poles"] == [(1.2, 0.1, 1.5), (0.0, 0.3, 0.5, 0.9)]
    mL, mD = out["masks"]
    assert np.array_equal(mL, (1 - cc) * (cb + (1 - cb) * ca))
    assert np.array_equal(mD, cc)
    assert np.array_equal(out["chi3"], cb * 0.1)
    sd, sb, mu = out["cond"]
    assert np.array_equal(sd, cc * 0.3) and not sb.any()
    v = np.ones((gs.nx, gs.ny))
    for cov, val in ((ca, 1.0), (cb, 1.0), (cc, 2.0)):                  # every object paints mu
        v = cov * val + (1.0 - cov) * v
    assert np.array_equal(mu, v)
    # a nonzero background chi3 is painted over everything, last
    out2 = G.rasterize_media([a, b], gs, mp.Medium(chi3=0.05))
    assert np.array_equal(out2["chi3"], np.full((gs.nx, gs.ny), 0.05))
    # no media anywhere: no maps
    out3 = G.rasterize_media([mp.Block(size=mp.Vector3(1, 1), material=mp.Medium(epsilon=4))], gs, mp.Medium())
    assert out3["poles"] == [] and out3["chi3"] is None and out3["cond"] is None


def test_media_in_3d_cells():
    """3D cells take D / B conductivity (painted like eps, as the reference's
    3D adapter does) and Lorentz / Drude poles (one mask per medium and
    pole, the cells that medium paints last); chi3 and mu raise by name."""
    import meep_torch_amd.mp as mp
    for med in (mp.Medium(epsilon=2, chi3=0.1), mp.Medium(epsilon=2, mu=2.0)):
        s3 = mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=3, device="cpu",
                           geometry=[mp.Block(size=mp.Vector3(0.5, 0.5, 0.5), material=med)])
        with pytest.raises(NotImplementedError, match="chi3 and mu"):
            s3.init_sim()
    with pytest.raises(NotImplementedError):
        mp.LorentzianSusceptibility(frequency=1.0, gamma=0.1, sigma_diag=mp.Vector3(1, 2, 1))
    drude = mp.DrudeSusceptibility(frequency=2.0, gamma=0.3, sigma=1.5)
    metal = mp.Medium(epsilon=2, D_conductivity=0.4, E_susceptibilities=[drude])
    lossy = mp.Medium(epsilon=3, B_conductivity=0.2)
    s3 = mp.Simulation(cell_size=mp.Vector3(1, 1, 1), resolution=10, dimensions=3, device="cpu",
                       dtype=torch.float64, boundary_layers=[mp.PML(0.2)],
                       geometry=[mp.Block(size=mp.Vector3(0.6, 0.6, 0.6), material=metal),
                                 mp.Block(center=mp.Vector3(0.2, 0, 0), size=mp.Vector3(0.2, 1, 1),
                                          material=lossy)])
    s3.init_sim()
    c = s3._coeffs
    assert c.pols is not None and c.pols["x"][0].numel() == 1
    # cell-centred maps recovered from the pole mask and eps: the Drude block minus the slab
    eps = s3._eps3d
    slab = eps == 3.0
    block = (eps == 2.0)
    assert int(block.sum()) > 0 and int(slab.sum()) > 0
    dt = c.dt
    v3 = drude.to_pole()
    from meep_torch_amd.core.coeffs import pole_recurrence
    r3 = pole_recurrence(v3, dt)[2]
    assert torch.equal(c.pols["z"][2][0], r3 * (0.5 * (block.double()[:, :, :-1] + block.double()[:, :, 1:])))
    sd = block.double() * 0.4
    sb = slab.double() * 0.2
    a = 0.5 * (sd[:-1] + sd[1:])
    assert torch.equal(c.damp["ex"][0], (1 - a * dt / 2) / (1 + a * dt / 2))
    b = 0.5 * (sb[:, :-1] + sb[:, 1:])
    b = 0.5 * (b[:, :, :-1] + b[:, :, 1:])
    assert torch.equal(c.damp["hx"][1], dt / (1 + b * dt / 2))


def test_eigenmode_source_port_mapping_and_refusals():
    """EigenModeSource / add_mode_monitor set-up (mp/sources.py:510-529,
    mp/simulation.py:1766-1789): a line normal to the region's smallest
    extent; a monitor on the source's line is the reflection port; options
    with no engine path raise by name."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.mp.sources import port_line
    gs_sim = mp.Simulation(cell_size=mp.Vector3(20, 8), resolution=10, device="cpu")
    gs = gs_sim.gs
    assert port_line(mp.Vector3(-6, 0), mp.Vector3(0, 4), gs) == ("x", 40, 20, 60)
    assert port_line(mp.Vector3(0, 2), mp.Vector3(4, 0), gs) == ("y", 60, 80, 120)
    src = mp.GaussianSource(0.15, fwidth=0.1)
    with pytest.raises(ValueError, match="1-based"):
        mp.EigenModeSource(src, eig_band=0)
    with pytest.raises(NotImplementedError, match="eig_kpoint"):
        mp.EigenModeSource(src, eig_kpoint=mp.Vector3(1, 0))
    with pytest.raises(NotImplementedError, match="eig_match_freq"):
        mp.EigenModeSource(src, eig_match_freq=False)
    sim = mp.Simulation(cell_size=mp.Vector3(20, 8), resolution=10, device="cpu",
                        sources=[mp.EigenModeSource(src, center=mp.Vector3(-6, 0), size=mp.Vector3(0, 4),
                                                    direction=1)],
                        boundary_layers=[mp.PML(1.0)])
    f = sim.add_mode_monitor(0.15, 0.04, 3, mp.FluxRegion(center=mp.Vector3(-6, 0), size=mp.Vector3(0, 4)))
    sim.add_mode_monitor(0.15, 0.04, 3, mp.FluxRegion(center=mp.Vector3(6, 0), size=mp.Vector3(0, 4),
                                                      direction=-1))
    assert sim._mode_ports == [("mode0", "x", 40, 20, 60, None), ("mode1", "x", 160, 20, 60, -1)]
    assert mp.get_flux_freqs(f) == pytest.approx([0.13, 0.15, 0.17])
    with pytest.raises(NotImplementedError, match="mode ports"):
        sim.get_eigenmode_coefficients(f, [1])
    with pytest.raises(NotImplementedError, match="step functions"):
        sim.run(mp.at_end(lambda s: None), n_steps=10)
    # without an EigenModeSource a mode monitor is a flux monitor
    plain = mp.Simulation(cell_size=mp.Vector3(20, 8), resolution=10, device="cpu",
                          sources=[mp.Source(src, component=mp.Ez, center=mp.Vector3(-6, 0))],
                          boundary_layers=[mp.PML(1.0)])
    m = plain.add_mode_monitor(0.15, 0.04, 3, mp.FluxRegion(center=mp.Vector3(6, 0), size=mp.Vector3(0, 4)))
    assert plain._monitors[m.name][0] == "flux"


def _shapes_sim(device):
    import meep_torch_amd.mp as mp
    geom = [mp.Cylinder(radius=0.7, center=mp.Vector3(-1.5, 0.3), material=mp.Medium(epsilon=6.0)),
            mp.Sphere(radius=0.5, center=mp.Vector3(1.2, -0.8), material=mp.Medium(epsilon=3.0)),
            mp.Prism(vertices=[mp.Vector3(0, 0.5), mp.Vector3(1.5, 1.2), mp.Vector3(0.8, 1.8)],
                     material=mp.Medium(epsilon=9.0)),
            mp.Block(size=mp.Vector3(1.2, 0.4), center=mp.Vector3(-0.2, -1.0), e1=mp.Vector3(1, 1),
                     e2=mp.Vector3(-1, 1), material=mp.Medium(epsilon=4.0)),
            mp.Ellipsoid(size=mp.Vector3(1.6, 0.8), center=mp.Vector3(2.0, 1.0), material=mp.Medium(epsilon=2.5))]
    sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10, geometry=geom,
                        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=mp.Ez,
                                           center=mp.Vector3(-2.4, 0.1))],
                        boundary_layers=[mp.PML(0.5)], dtype=torch.float64, device=device)
    fl = sim.add_flux(0.8, 0.4, 3, mp.FluxRegion(center=mp.Vector3(2.6, 0), size=mp.Vector3(0, 3)))
    return sim, fl


def test_shapes_eps_matches_reference():
    """Cylinder, Sphere, Prism, rotated Block and Ellipsoid (mp/geometry.py:
    95-240 of the reference) rasterised by the facade: the eps map equals
    the reference Simulation's (golden shapes_f64) to 1e-12."""
    import meep_torch_amd.mp as mp
    from conftest import golden
    sim, _ = _shapes_sim("cpu")
    eps = np.asarray(sim.get_array(component=mp.Dielectric))
    ref = golden("shapes_f64")["eps"]
    assert eps.shape == ref.shape
    assert np.abs(eps - ref).max() <= 1e-12 * np.abs(ref).max()
    assert len(np.unique(np.round(ref, 6))) > 6     # every shape leaves its own values
