"""Simulation facade end to end on the HIP engine: the GOLDEN_PLAIN scene
(100x60 TM guide, eps 12, 10-cell PML, 800 steps, 3 flux frequencies) against
the reference's own run, in FP32 (fields <= 1e-5 rel-L2, spectra <= 1e-4) and
FP64 (<= 1e-10); consecutive runs integrate to the accumulated clock like one
continuing run; DFT field arrays match a run_tm DFT monitor; step functions,
stop conditions and flux-data loading follow the reference's run protocol."""
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


# The reference's flux values for the same scene run as
# run(at_every(5.0, hook), until=40) (meep_torch tests/mp_parity/test_meep_flow.py:53-54)
GOLDEN_STEPPED = [2.5635746995635057e-06, 6.786305104894549e-06, 4.077624008546449e-06]


def test_stepped_run_matches_reference_golden():
    """Hooks fire at t = 0, 5, ..., 40 (meep_torch tests/mp_parity/
    test_meep_flow.py:111-118); the run is split into 100-step engine calls
    at the hook times, which regroups the monitor sums only: fluxes within
    1e-12 of the reference's stepped values, fields equal to the plain run."""
    mp, sim, fl = _plain(torch.float64)
    hooks = []
    sim.run(mp.at_every(5.0, lambda s: hooks.append(s.round_time())), until=40)
    assert hooks == [0.0, 5.0, 10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0]
    got = np.array(mp.get_fluxes(fl))
    print(f"\n[golden_stepped] flux rel {_rel(got, np.array(GOLDEN_STEPPED)):.2e}")
    assert np.allclose(got, GOLDEN_STEPPED, rtol=1e-12, atol=0)
    _, plain, _ = _plain(torch.float64)
    plain.run(until=40)
    assert np.array_equal(sim.get_array(component=mp.Ez), plain.get_array(component=mp.Ez))


def test_stop_condition_steps_one_at_a_time():
    """A stop condition is fed after every step (meep_torch tests/mp_parity/
    test_meep_flow.py:98-108: seven calls, then stop); the fields equal a
    plain 7-step run and get_field_point reads the cell holding the point."""
    mp, sim, _ = _plain(torch.float64)
    seen = []
    sim.run(until=[lambda s: seen.append(s.round_time()) or len(seen) >= 7])
    dt = sim.grid.dt
    assert sim._t_step == 7 and np.allclose(seen, [dt * i for i in range(1, 8)], rtol=0, atol=1e-12)
    _, ref, _ = _plain(torch.float64)
    ref.run(n_steps=7)
    ez = ref.get_array(component=mp.Ez)
    assert np.array_equal(sim.get_array(component=mp.Ez), ez)
    # point (-3, 0.2) of the 10 x 6 cell at resolution 10 -> cell (20, 32)
    assert sim.get_field_point(mp.Ez, mp.Vector3(-3, 0.2)) == float(ez[20, 32])
    eps = sim.get_epsilon()
    assert sim.get_field_point("Dz", mp.Vector3(-3, 0.2)) == float(eps[20, 32]) * float(ez[20, 32])


def test_stop_when_fields_decayed_after_sources():
    """until_after_sources=stop_when_fields_decayed(...): conditions count only
    after the source cutoff.  A plain run of the same length, with a per-step
    hook recording the probe, replays the rule offline and must stop at the
    same step."""
    mp, sim, _ = _plain(torch.float64)
    pt = mp.Vector3(3, 0)
    limit = lambda s: s.round_time() > 400.0     # bounds the test; must not be what stops it
    sim.run(until_after_sources=[mp.stop_when_fields_decayed(5.0, mp.Ez, pt, 1e-2), limit])
    n_stop, cutoff, dt = sim._t_step, sim._source_cutoff(), sim.grid.dt
    assert cutoff < n_stop * dt < 400.0
    _, ref, _ = _plain(torch.float64)
    probe = []
    ref.run(lambda s: probe.append(s.get_field_point(mp.Ez, pt)), n_steps=n_stop)
    cond = mp.stop_when_fields_decayed(5.0, mp.Ez, pt, 1e-2)

    class Replay:
        grid = ref.grid

        def __init__(self):
            self.i = 0

        def round_time(self):
            return self.i * dt

        def get_field_point(self, c, p):
            return probe[self.i - 1]

    r = Replay()
    stop = None
    for i in range(1, n_stop + 1):
        r.i = i
        if cond(r) and i * dt >= cutoff - 1e-9:
            stop = i
            break
    assert stop == n_stop


def test_flux_data_roundtrip():
    """load_minus_flux_data of a run's own data, then the same run: every
    accumulator is -a + a = 0 exactly, so the flux is exactly zero;
    load_flux_data instead doubles E and H (exact), so the flux is exactly 4x."""
    mp, a, fa = _plain(torch.float64)
    a.run(until=40)
    data = a.get_flux_data(fa)
    assert data.E.shape == (3, 40) and data.H.shape == (3, 40)
    _, b, fb = _plain(torch.float64)
    b.load_minus_flux_data(fb, data)
    b.run(until=40)
    assert mp.get_fluxes(fb) == [0.0, 0.0, 0.0]
    _, c, fc = _plain(torch.float64)
    c.load_flux_data(fc, data)
    c.run(until=40)
    assert mp.get_fluxes(fc) == [4.0 * v for v in mp.get_fluxes(fa)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_media_blocks_equal_direct_run(dtype):
    """A TM cell with a Lorentz block (eps 2) and a Drude + chi3 + D-conductivity
    block, both on cell boundaries so their masks are exactly 0 / 1, against
    run_tm on hand-built coefficients: fields and fluxes identical."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    L = mp.LorentzianSusceptibility(frequency=1.2, gamma=0.1, sigma=1.5)
    D = mp.DrudeSusceptibility(frequency=0.9, gamma=0.3, sigma=0.5)
    sim = mp.Simulation(
        cell_size=mp.Vector3(10, 6), resolution=10,
        geometry=[mp.Block(center=mp.Vector3(-2.5, 0), size=mp.Vector3(1, mp.inf),
                           material=mp.Medium(epsilon=2, E_susceptibilities=[L])),
                  mp.Block(center=mp.Vector3(1.5, 0), size=mp.Vector3(1, 2),
                           material=mp.Medium(E_susceptibilities=[D], chi3=0.05, D_conductivity=0.2))],
        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.5), component=mp.Ez,
                           center=mp.Vector3(-4, 0), size=mp.Vector3(0, 2))],
        boundary_layers=[mp.PML(1.0)], dtype=dtype, device="cuda")
    fl = sim.add_flux(0.8, 0.4, 3, mp.FluxRegion(center=mp.Vector3(4, 0), size=mp.Vector3(0, 4)))
    sim.run(until=20)
    # by hand: x corner [2, 3) -> i 20..29 (all y); x [6, 7), y [2, 4) -> i 60..69, j 20..39
    g = sim.grid
    z = lambda v=0.0: torch.full((100, 60), v, dtype=dtype, device="cuda")
    eps, mL, mD, chi3, sd = z(1.0), z(), z(), z(), z()
    eps[20:30, :] = 2.0
    mL[20:30, :] = 1.0
    mD[60:70, 20:40] = 1.0
    chi3[60:70, 20:40] = 0.05
    sd[60:70, 20:40] = 0.2
    m = build_tm_coeffs(g, eps_inf=eps, npml=((10, 10), (10, 10)), poles=[L.to_pole(), D.to_pole()],
                        pole_masks=[mL, mD], chi3=chi3, sigma_d=sd, sigma_b=z(), mu=z(1.0))
    src = sim._make_sources()[0]
    freqs = np.linspace(0.8 - 0.4 / 2.0, 0.8 + 0.4 / 2.0, 3).tolist()     # meep's fcen +- df/2 grid
    mon = FluxMonitorTM(g, freqs, x_normal=True, idx=90, lo=10, hi=50)
    st, (acc,) = run_tm(init_tm_state(g, n_poles=2), m, sim._t_step, source=src, monitors=[mon])
    assert sim._t_step == 400
    assert np.array_equal(sim.get_array(component=mp.Ez), st.Ez.cpu().numpy())
    assert np.array_equal(np.asarray(mp.get_fluxes(fl)), mon.flux(acc, g.dx).cpu().numpy())
    assert float(st.Ez.abs().max()) > 0 and float(st.P_cur.abs().max()) > 0


def test_shapes_run_matches_reference():
    """The facade cell with a Cylinder, Sphere, Prism, rotated Block and
    Ellipsoid, run to t = 12 on the HIP engine: Ez and the flux spectrum
    equal the reference Simulation's (golden shapes_f64) within FP64
    rounding."""
    import meep_torch_amd.mp as mp
    from conftest import golden
    from test_mp_facade_cpu import _shapes_sim
    G = golden("shapes_f64")
    sim, fl = _shapes_sim("cuda")
    sim.run(until=12)
    ez = np.asarray(sim.get_array(component=mp.Ez))
    assert np.linalg.norm(ez - G["Ez"]) <= 1e-10 * np.linalg.norm(G["Ez"])
    fx = np.asarray(mp.get_fluxes(fl))
    assert np.abs(fx - G["flux"]).max() <= 1e-9 * np.abs(G["flux"]).max()


@pytest.mark.parametrize("hook", [False, True])
def test_several_sources_match_reference(hook):
    """Two sources in one facade run (a point and a line current with their
    own pulses and amplitudes): the engine injects sum_i amp_i(t) mask_i each
    step from a per-step weight stack (mp/cpp/core.py:584-589), and with a
    step function splitting the run the spectra keep one running sum across
    the engine calls, as the reference's per-step loop does.  Ez, the flux
    spectrum and the probed values equal the reference Simulation's (golden
    multisrc_f64) within FP64 rounding."""
    import meep_torch_amd.mp as mp
    G = golden("multisrc_f64")
    sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10,
                        geometry=[mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))],
                        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=mp.Ez,
                                           center=mp.Vector3(-2.0, 0.0)),
                                 mp.Source(mp.GaussianSource(1.1, fwidth=0.5), component=mp.Ez,
                                           center=mp.Vector3(-1.4, 0.5), size=mp.Vector3(0, 0.8), amplitude=-0.7)],
                        boundary_layers=[mp.PML(0.5)], dtype=torch.float64, device="cuda")
    fl = sim.add_flux(0.9, 0.5, 4, mp.FluxRegion(center=mp.Vector3(2.2, 0), size=mp.Vector3(0, 3)))
    seen = []
    if hook:
        sim.run(mp.at_every(1.5, lambda s: seen.append(float(s.get_field_point(mp.Ez, mp.Vector3(1.0, 0))))),
                until=10)
    else:
        sim.run(until=10)
    tag = "hook" if hook else "plain"
    ez = np.asarray(sim.get_array(component=mp.Ez))
    assert np.linalg.norm(ez - G[f"{tag}_Ez"]) <= 1e-12 * np.linalg.norm(G[f"{tag}_Ez"])
    fx = np.asarray(mp.get_fluxes(fl))
    assert np.abs(fx - G[f"{tag}_flux"]).max() <= 1e-10 * np.abs(G[f"{tag}_flux"]).max()
    if hook:
        assert np.allclose(seen, G["hook_probe"], rtol=1e-10, atol=1e-16)


@pytest.mark.parametrize("pol", ["tm", "te"])
def test_decay_stop_in_window_sized_calls(pol, monkeypatch):
    """stop_when_fields_decayed compares only at the end of its dt window
    (mp/run_control.py:323-351 of the reference): the facade runs a whole
    window per engine call with a device probe on the watched component, and
    must stop at the same step, with the same fields and spectra, as the
    same condition fed after every step (a wrapping callable forces that
    path), in far fewer engine calls."""
    import meep_torch_amd.mp as mp
    from meep_torch_amd.core import engine2d
    calls = {"n": 0}
    real = engine2d.run_chunk

    def counting(*a, **k):
        calls["n"] += 1
        return real(*a, **k)

    monkeypatch.setattr(engine2d, "run_chunk", counting)
    comp = mp.Ez if pol == "tm" else mp.Hz

    def run(wrapped):
        sim = mp.Simulation(cell_size=mp.Vector3(8, 4), resolution=10,
                            geometry=[mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=9.0))],
                            sources=[mp.Source(mp.GaussianSource(0.7, fwidth=0.4), component=comp,
                                               center=mp.Vector3(-3.0, 0))],
                            boundary_layers=[mp.PML(0.8)], dtype=torch.float64, device="cuda")
        fl = sim.add_flux(0.7, 0.3, 3, mp.FluxRegion(center=mp.Vector3(3.0, 0), size=mp.Vector3(0, 3)))
        cond = mp.stop_when_fields_decayed(4.0, comp, mp.Vector3(3.0, 0), 1e-3)
        calls["n"] = 0
        sim.run(until_after_sources=(lambda s: cond(s)) if wrapped else cond)
        return sim.round_time(), np.asarray(sim.get_array(component=comp)), np.asarray(mp.get_fluxes(fl)), calls["n"]

    t_w, f_w, x_w, n_w = run(True)
    t_f, f_f, x_f, n_f = run(False)
    assert t_f == t_w
    assert np.array_equal(f_f, f_w) and np.array_equal(x_f, x_w)
    assert n_f * 20 < n_w, (n_f, n_w)


@pytest.mark.parametrize("pol", ["tm", "te"])
def test_absorber_layers_match_reference(pol):
    """mp.Absorber on x with a PML on y: the facade folds the matched
    absorber ramp (max of the x and y absorber_sigma ramps) into sigma_d and
    sigma_b (mp/simulation.py:539-550,615-627) and runs the conductive
    kernels; the field and the flux equal the reference Simulation's
    (golden sim_absorber_f64) within FP64 rounding."""
    import meep_torch_amd.mp as mp
    G = golden("sim_absorber_f64")
    comp = mp.Ez if pol == "tm" else mp.Hz
    sim = mp.Simulation(cell_size=mp.Vector3(6, 4), resolution=10,
                        geometry=[mp.Block(size=mp.Vector3(mp.inf, 0.6), material=mp.Medium(epsilon=8.0))],
                        sources=[mp.Source(mp.GaussianSource(0.8, fwidth=0.6), component=comp,
                                           center=mp.Vector3(-1.8, 0.0), size=mp.Vector3(0, 0.8))],
                        boundary_layers=[mp.Absorber(0.8, direction=mp.X), mp.PML(0.5, direction=mp.Y)],
                        dtype=torch.float64, device="cuda")
    fl = sim.add_flux(0.8, 0.5, 4, mp.FluxRegion(center=mp.Vector3(1.6, 0), size=mp.Vector3(0, 3)))
    sim.run(until=12)
    f = np.asarray(sim.get_array(component=comp))
    assert f.shape == G[f"{pol}_field"].shape
    assert np.linalg.norm(f - G[f"{pol}_field"]) <= 1e-12 * np.linalg.norm(G[f"{pol}_field"])
    fx = np.asarray(mp.get_fluxes(fl))
    assert np.abs(fx - G[f"{pol}_flux"]).max() <= 1e-10 * np.abs(G[f"{pol}_flux"]).max()
