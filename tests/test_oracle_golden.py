"""Pin the CPU oracle to the reference: golden vectors written by the reference
itself (tests/golden/make_golden.py) must be reproduced BIT-EXACTLY by the
oracle's restatement on this CPU, in FP32 and FP64.  Bit-exact equality means
a dropped term, a swapped index or a changed operation order fails here."""
import json
import math

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import yee2d, yee3d

torch.set_num_threads(8)
FREQS5 = [0.3, 0.4, 0.5, 0.6, 0.7]


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def _eq(a, b):
    a = a.detach() if isinstance(a, torch.Tensor) else _t(a)
    b = b.detach() if isinstance(b, torch.Tensor) else _t(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert a.dtype == b.dtype, (a.dtype, b.dtype)
    assert torch.equal(a, b), f"max |diff| {float((a - b).abs().max()):.3e}"


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_cpml_profiles_bitexact(tag, dtype):
    g = golden("pml")
    combos = json.loads(bytes(g["combos_json"]).decode())
    for ci, (n, npml, off) in enumerate(combos):
        npml = tuple(npml) if isinstance(npml, list) else npml
        b, a, k = yee2d.cpml_profiles(n, npml, off, 0.5 * 0.05, 0.05, dtype)
        _eq(b, g[f"{tag}_{ci}_b"])
        _eq(a, g[f"{tag}_{ci}_a"])
        _eq(k, g[f"{tag}_{ci}_k"])


def test_cpml_profile_closed_form():
    """Interior identity, negative a inside the layer, mirror symmetry."""
    b, a, k = yee2d.cpml_profiles(100, 10, 0.0, 0.01, 0.02)
    assert torch.all(b[20:80] == 1.0) and torch.all(a[20:80] == 0.0)
    assert torch.all(a[:5] < 0) and torch.equal(b, b.flip(0))
    # sigma at the outer node of a 10-cell layer: -(4 ln 1e-6)/(2*10*0.02)
    sig0 = -4 * math.log(1e-6) / (2 * 10 * 0.02)
    assert float(b[0]) == pytest.approx(math.exp(-sig0 * 0.01), rel=1e-12)


def _tm_small(tag, dtype, sname, ck):
    g = golden(f"tm_small_{tag}")
    eps = _t(g["eps"])
    c = yee2d.tm_coeffs(64, 48, 0.05, eps, 8, dtype=dtype)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 50, "lo": 10, "hi": 38, "freqs": FREQS5},
            {"type": "flux", "kind": "tm_y", "idx": 36, "lo": 10, "hi": 54, "freqs": FREQS5},
            {"type": "dft", "comp": "Ez", "freqs": [0.45, 0.55]},
            {"type": "dft", "comp": "Hx", "freqs": [0.5]}]
    amps = g[f"{sname}_amp"]
    s, acc = yee2d.run2d("tm", yee2d.tm_state(64, 48, dtype), c, 400, amps=amps,
                         mask=_t(g[f"{sname}_mask"]), monitors=mons, checkpoint_every=ck)
    return g, s, acc, c


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("sname", ["point", "line"])
@pytest.mark.parametrize("ck", [0, 150])
def test_tm_small_bitexact(tag, dtype, sname, ck):
    g, s, acc, c = _tm_small(tag, dtype, sname, ck)
    p = f"{sname}_ck{ck}"
    for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
        _eq(s[f], g[f"{p}_{f}"])
    for i in range(4):
        _eq(acc[i], g[f"{p}_acc{i}"])
    _eq(yee2d.flux_from_acc("tm_x", acc[0], 0.05), g[f"{p}_fluxx"])
    _eq(yee2d.flux_from_acc("tm_y", acc[1], 0.05), g[f"{p}_fluxy"])


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_inconsistent_initial_state(tag, dtype):
    """The first B half reads the given Ez, not Dz/eps (stepper_tm.py:87)."""
    g = golden(f"tm_small_{tag}")
    c = yee2d.tm_coeffs(64, 48, 0.05, _t(g["eps"]), 8, dtype=dtype)
    s = yee2d.tm_state(64, 48, dtype)
    G = _t(g["blob_Ez0"])
    s["Ez"], s["Dz"] = G.clone(), 1.5 * G
    s, _ = yee2d.run2d("tm", s, c, 50)
    for f in ("Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"):
        _eq(s[f], g[f"blob_{f}"])


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("ck", [0, 150])
def test_te_small_bitexact(tag, dtype, ck):
    g = golden(f"te_small_{tag}")
    c = yee2d.te_coeffs(64, 48, 0.05, _t(g["eps"]), 8, dtype=dtype)
    mons = [{"type": "flux", "kind": "te_x", "idx": 50, "lo": 10, "hi": 38, "freqs": FREQS5},
            {"type": "flux", "kind": "te_y", "idx": 36, "lo": 10, "hi": 54, "freqs": FREQS5},
            {"type": "dft", "comp": "Hz", "freqs": [0.45, 0.55]},
            {"type": "dft", "comp": "Ey", "freqs": [0.5]}]
    s, acc = yee2d.run2d("te", yee2d.te_state(64, 48, dtype), c, 400, amps=g["amp"],
                         mask=_t(g["mask"]), monitors=mons, checkpoint_every=ck)
    p = f"ck{ck}"
    for f in ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"):
        _eq(s[f], g[f"{p}_{f}"])
    for i in range(4):
        _eq(acc[i], g[f"{p}_acc{i}"])
    _eq(yee2d.flux_from_acc("te_x", acc[0], 0.05), g[f"{p}_fluxx"])
    _eq(yee2d.flux_from_acc("te_y", acc[1], 0.05), g[f"{p}_fluxy"])


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_c1_full_bitexact(tag, dtype):
    """C1 at full size: 256x256, 2000 steps, 50-frequency flux line."""
    g = golden(f"c1_{tag}")
    c = yee2d.tm_coeffs(256, 256, 0.05, _t(g["eps"]), 16, dtype=dtype)
    amp = yee2d.gaussian_pulse(0.15, 0.1)
    amps = [amp((t + 0.5) * c.dt) for t in range(2000)]
    mask = yee2d.line_mask(256, 256, 24, 128, 129, 0.05, dtype)
    mon = {"type": "flux", "kind": "tm_x", "idx": 232, "lo": 96, "hi": 160,
           "freqs": np.linspace(0.1, 0.2, 50).tolist()}
    s, (acc,) = yee2d.run2d("tm", yee2d.tm_state(256, 256, dtype), c, 2000, amps=amps,
                            mask=mask, monitors=[mon])
    for f in ("Ez", "Hx", "Hy"):
        _eq(s[f], g[f])
    _eq(acc, g["acc"])
    _eq(yee2d.flux_from_acc("tm_x", acc, 0.05), g["flux"])


def test_tm_yee_invariant_pec_box():
    """Closed-form gate (tests/analytic/test_energy.py:44-61): the discrete
    leapfrog energy is conserved to FP64 round-off in a lossless PEC box."""
    n, dx = 100, 0.02           # blob is exp(-50) at the walls: a true PEC start
    c = yee2d.tm_coeffs(n, n, dx, torch.ones(n, n, dtype=torch.float64), 0,
                        dtype=torch.float64)
    s = yee2d.tm_state(n, n, torch.float64)
    x = torch.arange(n, dtype=torch.float64)
    blob = torch.exp(-((x[:, None] - 50) ** 2 + (x[None, :] - 50) ** 2) / (2 * 5.0 ** 2))
    s["Ez"], s["Dz"] = blob.clone(), blob.clone()
    inv = []
    for i in range(600):
        prev = s
        s = yee2d.tm_step(s, c)
        if i % 50 == 0:
            inv.append(yee2d.tm_yee_invariant(prev, s, c.eps, dx))
    inv = np.array(inv)
    assert (inv.max() - inv.min()) / inv.mean() < 1e-12


def test_te_tm_vacuum_duality():
    """tests/analytic/test_te.py:21-39: in vacuum TE Hz evolves exactly like TM Ez."""
    n = 80
    x = torch.arange(n, dtype=torch.float64)
    G = torch.exp(-((x[:, None] - 40.0) ** 2 + (x[None, :] - 40.0) ** 2) / (2 * 4.0 ** 2))
    ctm = yee2d.tm_coeffs(n, n, 0.02, torch.ones(n, n, dtype=torch.float64), 0, dtype=torch.float64)
    cte = yee2d.te_coeffs(n, n, 0.02, torch.ones(n, n, dtype=torch.float64), 0, dtype=torch.float64)
    stm = yee2d.tm_state(n, n, torch.float64)
    stm["Ez"], stm["Dz"] = G.clone(), G.clone()
    ste = yee2d.te_state(n, n, torch.float64)
    ste["Hz"] = G[:-1, :-1].clone()
    for _ in range(20):
        stm = yee2d.tm_step(stm, ctm)
        ste = yee2d.te_step(ste, cte)
    w = slice(20, 60)
    assert torch.allclose(ste["Hz"][w, w], stm["Ez"][w, w], atol=1e-12, rtol=0)


def test_dft_closed_form():
    """DFT of a pure cosine field at its own frequency: the accumulated phasor
    approaches (N dt / 2) e^{0} (tests/examples/test_w3_dft_fields.py:19-37)."""
    f0, dt, N = 0.5, 0.025, 4000
    acc = None
    for t in range(N):
        F = torch.full((1, 1), math.cos(2 * math.pi * f0 * (t + 1.0) * dt), dtype=torch.float64)
        smp = yee2d.dft_field_sample(F, "Ez", [f0], dt, t, torch.float64)
        acc = smp if acc is None else acc + smp
    re, im = float(acc[0, 0, 0, 0]), float(acc[0, 1, 0, 0])
    assert re == pytest.approx(N * dt / 2, rel=1e-3) and abs(im) < 1e-2


def _d3(tag, dtype, sname, lean=False):
    g = golden(f"d3_small_{tag}")
    nx, ny, nz, npml, dx, n = 32, 28, 24, 5, 0.05, 200
    c = yee3d.coeffs3d(nx, ny, nz, dx, _t(g["eps"]), npml, dtype)
    amp = yee2d.gaussian_pulse(1.0, 2.0, cutoff=2.0)
    amps = [amp((t + 0.5) * c["dt"]) for t in range(n)]
    if sname == "point":
        src = ("Ez", 8, ny // 2, nz // 2)
    else:
        mask = torch.zeros(nx, ny - 1, nz, dtype=torch.float64)
        mask[8, npml:min(nx, ny, nz) - npml, npml:min(nx, ny, nz) - npml] = 1.0
        src = ("Ey", mask)
    freqs = [0.8, 1.0, 1.2, 1.4]
    mons = [{"type": "flux", "normal": "x", "idx": 24, "freqs": freqs, "ap": npml + 1}]
    if not lean:
        mons += [{"type": "flux", "normal": "z", "idx": 14, "freqs": freqs, "ap": npml + 1},
                 {"type": "dft", "comp": "Ez", "plane": (2, 12), "freq": 1.1},
                 {"type": "dft", "comp": "Hy", "plane": (0, 20), "freq": 1.0}]
    st, acc = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype), c, n, amps, src, mons, lean=lean)
    return g, st, acc


@pytest.mark.parametrize("tag,dtype,sname", [("f32", torch.float32, "point"),
                                             ("f64", torch.float64, "point"),
                                             ("f64", torch.float64, "plane")])
def test_3d_small_bitexact(tag, dtype, sname):
    g, st, acc = _d3(tag, dtype, sname)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "psi_Ez_x"):
        _eq(st[f], g[f"{sname}_{f}"])
    for i in range(4):
        _eq(acc[i], g[f"{sname}_acc{i}"])
    _eq(yee3d.flux_plane(acc[0], 0.05), g[f"{sname}_flux0"])
    _eq(yee3d.flux_plane(acc[1], 0.05), g[f"{sname}_flux1"])


def test_3d_lean_bitexact():
    g, st, acc = _d3("f32", torch.float32, "point", lean=True)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        _eq(st[f], g[f"lean_{f}"])
    _eq(acc[0], g["lean_acc0"])
