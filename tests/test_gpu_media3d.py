"""3D material conductivity and Lorentz / Drude ADE poles
(meep_torch/core/stepper_3d.py:217-266,295-301,318-329) on the HIP two-pass
kernels against the oracle, whose 3D media are pinned against the 2D TM
oracle by tests/test_oracle_media3d_cpu.py.

Checked: a CPML run with sigma_d, sigma_b and two poles (one Lorentz, one
Drude) under a point and a mask source, fields, pole stacks, CPML memory and
flux / DFT spectra (FP64 to 1e-12, FP32 within the 1e-5 / 1e-4 bars);
chunked runs that carry the pole stacks across calls; the single-step
wrapper from a random state; each medium alone."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, NZ, NPML, DX, N = 32, 28, 24, 5, 0.05, 200
POLES = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]
FREQS = [0.8, 1.0, 1.2]


def _maps(sigma=True, poles=True):
    """Cell-centred eps, sigma_d, sigma_b and two pole masks (FP64, CPU);
    the media sit inside the CPML shells' inner edge."""
    g = torch.Generator().manual_seed(11)
    r = lambda: torch.rand(NX, NY, NZ, generator=g, dtype=torch.float64)
    inner = torch.zeros(NX, NY, NZ, dtype=torch.float64)
    inner[NPML + 2:NX - NPML - 2, NPML + 2:NY - NPML - 2, NPML + 2:NZ - NPML - 2] = 1.0
    eps = 1.0 + 2.0 * r() * inner
    out = {"eps": eps}
    if sigma:
        out.update(sigma_d=3.0 * r() * inner, sigma_b=1.5 * r() * inner)
    if poles:
        half = (r() > 0.5).double() * inner
        out.update(poles=POLES, pole_masks=[half, inner - half])
    return out


def _source(kind, dtype):
    if kind == "point":
        return ("Ez", 9, NY // 2, NZ // 2)
    m = torch.zeros(NX, NY - 1, NZ, dtype=dtype)
    m[9, NPML:NY - NPML, NPML:NZ - NPML] = 1.0 / DX
    return ("Ey", m)


def _amps(dt, n):
    from oracle import yee2d
    amp = yee2d.gaussian_pulse(1.0, 2.0, cutoff=2.0)
    return [amp((t + 0.5) * dt) for t in range(n)]


def _hip(mp, dtype, kind, n=N, chunks=1):
    from meep_torch_amd.core.monitors_3d import DFTFieldsMonitor3D, FluxMonitor3D, run_3d
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    kw = {k: (v.to(dtype).cuda() if torch.is_tensor(v) else v) for k, v in mp.items() if k != "pole_masks"}
    if "pole_masks" in mp:
        kw["pole_masks"] = [m.to(dtype).cuda() for m in mp["pole_masks"]]
    c = build_3d_coeffs(NX, NY, NZ, DX, npml=NPML, dtype=dtype, device="cuda", **kw)
    amps = _amps(c.dt, n)
    src_t = _source(kind, dtype)
    if len(src_t) == 2:
        src_t = (src_t[0], src_t[1].cuda())

    class Src:
        src = src_t

        def __init__(self, off):
            self.off = off

        def amplitude(self, t):
            return amps[t + self.off]

    mons = [FluxMonitor3D(c.dt, FREQS, normal="x", idx=24, ap=NPML + 1, dtype=dtype),
            FluxMonitor3D(c.dt, FREQS, normal="z", idx=14, ap=NPML + 1, dtype=dtype),
            DFTFieldsMonitor3D(c.dt, [1.1], "Ez", plane=(2, 12), dtype=dtype)]
    st = init_3d_state(NX, NY, NZ, npml=NPML, n_poles=len(mp.get("poles", ())), dtype=dtype, device="cuda")
    step, acc = n // chunks, None
    for q in range(chunks):
        st, a = run_3d(st, c, step, source=Src(q * step), monitors=mons)
        acc = a if acc is None else [x + y for x, y in zip(acc, a)]
    return st, acc, mons


def _oracle(mp, dtype, kind, n=N):
    from oracle import yee3d
    torch.set_num_threads(8)
    kw = {k: (v.to(dtype) if torch.is_tensor(v) else v) for k, v in mp.items() if k not in ("eps", "pole_masks")}
    if "pole_masks" in mp:
        kw["pole_masks"] = [m.to(dtype) for m in mp["pole_masks"]]
    c = yee3d.coeffs3d(NX, NY, NZ, DX, mp["eps"], NPML, dtype, **kw)
    mons = [{"type": "flux", "normal": "x", "idx": 24, "freqs": FREQS, "ap": NPML + 1},
            {"type": "flux", "normal": "z", "idx": 14, "freqs": FREQS, "ap": NPML + 1},
            {"type": "dft", "comp": "Ez", "plane": (2, 12), "freq": 1.1}]
    st = yee3d.state3d(NX, NY, NZ, dtype, n_poles=len(mp.get("poles", ())))
    return yee3d.run3d(st, c, n, _amps(c["dt"], n), _source(kind, dtype), mons)


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / b.norm())


FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "psi_Ez_x", "psi_Hy_z", "Px_cur", "Py_prev", "Pz_cur")


@pytest.mark.parametrize("kind", ["point", "mask"])
@pytest.mark.parametrize("dtype,tf,ts", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)],
                         ids=["f64", "f32"])
def test_media3d_run_matches_oracle(kind, dtype, tf, ts):
    mp = _maps()
    st, acc, mons = _hip(mp, dtype, kind)
    s, oacc = _oracle(mp, dtype, kind)
    errs = {f: rel(getattr(st, f), s[f]) for f in FIELDS}
    errs.update({f"acc{i}": rel(a, b) for i, (a, b) in enumerate(zip(acc, oacc))})
    print(f"\n[media3d {kind} {dtype}] " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for f in FIELDS:
        assert errs[f] <= tf, (f, errs[f])
    for i in range(3):
        assert errs[f"acc{i}"] <= ts
    assert float(s["Px_cur"].abs().max()) > 0
    from oracle import yee3d
    for i in range(2):
        assert rel(mons[i].flux(acc[i], DX), yee3d.flux_plane(oacc[i], DX)) <= ts


@pytest.mark.parametrize("sigma,poles", [(True, False), (False, True)], ids=["sigma", "poles"])
def test_each_medium_alone(sigma, poles):
    mp = _maps(sigma=sigma, poles=poles)
    st, acc, _ = _hip(mp, torch.float64, "point")
    s, oacc = _oracle(mp, torch.float64, "point")
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert rel(getattr(st, f), s[f]) <= 1e-12, f
    bare, _, _ = _hip({"eps": mp["eps"]}, torch.float64, "point")
    assert rel(bare.Ez, s["Ez"]) > 1e-3


def test_chunked_runs_carry_the_pole_stacks():
    """Four calls of 50 steps leave the state of one call of 200, fields and
    pole stacks bit for bit (each run_3d call starts its monitors' clock at
    its own first step, as monitors_3d.py:192-210 does, so only the state is
    compared)."""
    mp = _maps()
    a, _, _ = _hip(mp, torch.float64, "mask")
    b, _, _ = _hip(mp, torch.float64, "mask", chunks=4)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Px_cur", "Px_prev", "Pz_cur", "Pz_prev", "psi_Hy_z"):
        assert torch.equal(getattr(a, f), getattr(b, f)), f


def test_step_wrapper_with_media_from_a_random_state():
    """step_3d with media from a random field and pole state (psi zero
    outside the shells, as every run keeps it) against the oracle's step."""
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state, state_names, step_3d
    from oracle import yee3d
    dt = torch.float64
    mp = _maps()
    gen = torch.Generator().manual_seed(5)
    st = init_3d_state(NX, NY, NZ, npml=NPML, n_poles=2, dtype=dt, device="cuda")
    rnd = {f: torch.randn(getattr(st, f).shape, generator=gen, dtype=dt)
           for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Px_cur", "Px_prev", "Py_cur", "Py_prev",
                     "Pz_cur", "Pz_prev")}
    st = st._replace(**{f: v.cuda() for f, v in rnd.items()})
    kw = {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in mp.items() if k != "pole_masks"}
    c = build_3d_coeffs(NX, NY, NZ, DX, npml=NPML, dtype=dt, device="cuda",
                        pole_masks=[m.cuda() for m in mp["pole_masks"]], **kw)
    out = step_3d(st, c, 0.7, ("Ez", 9, NY // 2, NZ // 2))
    oc = yee3d.coeffs3d(NX, NY, NZ, DX, mp["eps"], NPML, dt, sigma_d=mp["sigma_d"], sigma_b=mp["sigma_b"],
                        poles=POLES, pole_masks=mp["pole_masks"])
    ost = {f: getattr(st, f).cpu() for f in state_names(st)}
    ref = yee3d.step3d(ost, oc, 0.7, ("Ez", 9, NY // 2, NZ // 2))
    for f in state_names(st):
        assert rel(getattr(out, f), ref[f]) <= 1e-13 or not ref[f].any(), f
    assert torch.equal(st.Ex.cpu(), rnd["Ex"])          # the input state is left alone


def test_media_refuse_the_lean_stepper():
    from meep_torch_amd.core.engine3d import run_fields
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state, state_names
    mp = _maps(poles=False)
    kw = {k: v.cuda() for k, v in mp.items()}
    c = build_3d_coeffs(NX, NY, NZ, DX, npml=NPML, dtype=torch.float64, device="cuda", **kw)
    st = init_3d_state(NX, NY, NZ, npml=NPML, dtype=torch.float64, device="cuda")
    work = {f: getattr(st, f).clone() for f in state_names(st)}
    with pytest.raises(NotImplementedError, match="lean"):
        run_fields(work, c, 3, None, [], lean=True)


def test_lean_poles_match_the_reference_run():
    """The lean stepper with two ADE poles on masks (stepper_3d_lean.py:143-165,
    207-231) against the reference's own run (lean_poles_3d fixture): fields
    and the P(n+1) / P(n) stacks bit for bit in FP32 (the fused recurrence
    fma(c3, E, fma(c1, P, c2 P_prev)) and E + ((curl dt + sum P) - sum P') /
    eps_c), the flux spectrum within 1e-4."""
    import numpy as np
    from conftest import golden
    from meep_torch_amd.core.monitors_3d import FluxMonitor3D, PointSource3D
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d_lean import build_lean, run_3d_lean
    g = golden("lean_poles_3d_f32")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    nx, ny, nz, npml, n = 22, 20, 18, 4, 120
    c = build_lean(nx, ny, nz, 0.05, t(g["eps"]).cuda(), npml, poles=POLES,
                   pole_masks=[t(g["m1"]).cuda(), t(g["m2"]).cuda()], dtype=torch.float32, device="cuda")
    src = PointSource3D(c.dt, "Ez", 7, ny // 2, nz // 2, GaussianPulse(1.0, 2.0, cutoff=2.0))
    mon = FluxMonitor3D(c.dt, FREQS, normal="x", idx=15, ap=npml + 1, dtype=torch.float32)
    view, (acc,) = run_3d_lean(c, n, source=src, monitors=[mon])
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert torch.equal(getattr(view, f).cpu(), t(g[f])), f
    for key in "xyz":
        assert torch.equal(c.P[key][0].cpu(), t(g[f"P{key}_new"])), key
        assert torch.equal(c.P[key][1].cpu(), t(g[f"P{key}_old"])), key
    # the FP32 spectrum: the staged monitor flush sums the steps in its own
    # order, so within the 1e-4 spectrum bar rather than bit for bit
    ref = t(g["acc"])
    assert float((acc.cpu() - ref).norm() / ref.norm()) <= 1e-4
    assert float(view.Ez.abs().max()) > 0 and float(c.P["z"][0].abs().max()) > 0
