"""Pins the 3D oracle's media (conductivity damping pairs and ADE poles,
meep_torch/core/stepper_3d.py:217-266,295-301) against the 2D TM oracle,
whose media are pinned by tests/test_oracle_media_cpu.py.

Between PEC walls, with no CPML, a 3D field made of one E component E_v along
an axis v plus the two H components across it, with every field, eps, sigma
and pole mask invariant along v, is an exact mode of the 3D scheme: E_v is
normal to the two walls across v, and every difference along v vanishes.  Its
update is the 2D TM update on the plane of the other two axes (a, b), taken in
cyclic order (x, y) for v = z, (y, z) for v = x, (z, x) for v = y, so that
curl signs carry over: E_v <-> Ez, H_a <-> Hx, H_b <-> Hy.  The two schemes
differ in form (the 3D one folds D = eps E + sum P into the E update), so the
comparison is to FP64 rounding.  Random, non-uniform media maps make an
averaging axis, a sign or a dropped term show."""
import math

import pytest
import torch

from oracle import yee2d, yee3d

NA, NB, L, DX, N = 14, 12, 3, 0.05, 160
POLES = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]


def _maps(seed):
    g = torch.Generator().manual_seed(seed)
    r = lambda: torch.rand(NA, NB, generator=g, dtype=torch.float64)
    m = (r() > 0.5).double()
    return {"eps": 1.0 + 2.0 * r(), "sd": 3.0 * r(), "sb": 2.0 * r(), "m0": m, "m1": 1.0 - m,
            "Ez": torch.randn(NA, NB, generator=g, dtype=torch.float64),
            "Hx": torch.randn(NA, NB - 1, generator=g, dtype=torch.float64),
            "Hy": torch.randn(NA - 1, NB, generator=g, dtype=torch.float64)}


def _lift(v, a2):
    """Cell-centred 2D map (NA, NB) -> 3D map invariant along axis v."""
    if v == 2:
        return a2[:, :, None].expand(NA, NB, L).contiguous()
    if v == 0:
        return a2[None].expand(L, NA, NB).contiguous()
    return a2.T[:, None, :].expand(NB, L, NA).contiguous()


def _dims(v):
    return {2: (NA, NB, L), 0: (L, NA, NB), 1: (NB, L, NA)}[v]


def _names(v):
    """(E_v, H_a, H_b) of the 3D state for invariant axis v."""
    return {2: ("Ez", "Hx", "Hy"), 0: ("Ex", "Hy", "Hz"), 1: ("Ey", "Hz", "Hx")}[v]


def _plane(v, t):
    """3D component (or pole stack) -> its 2D plane at index 0 along v."""
    if v == 2:
        return t[..., 0]
    if v == 0:
        return t[..., 0, :, :]
    return t[..., :, 0, :].transpose(-1, -2)


def _put(v, t3, p2):
    """Write the 2D plane p2 at every index along v of the 3D component t3."""
    if v == 2:
        t3[...] = p2[:, :, None]
    elif v == 0:
        t3[...] = p2[None]
    else:
        t3[...] = p2.T[:, None, :]


def _run(v, mp, sigma=True, poles=True):
    nx, ny, nz = _dims(v)
    dt = 0.5 * DX / math.sqrt(3.0)
    kw2, kw3 = {}, {}
    if sigma:
        kw2.update(sigma_d=mp["sd"], sigma_b=mp["sb"])
        kw3.update(sigma_d=_lift(v, mp["sd"]), sigma_b=_lift(v, mp["sb"]))
    if poles:
        kw2.update(poles=POLES, pole_masks=[mp["m0"], mp["m1"]])
        kw3.update(poles=POLES, pole_masks=[_lift(v, mp["m0"]), _lift(v, mp["m1"])])
    c2 = yee2d.tm_coeffs(NA, NB, DX, mp["eps"], 0, courant=dt / DX, dtype=torch.float64, **kw2)
    c3 = yee3d.coeffs3d(nx, ny, nz, DX, _lift(v, mp["eps"]), 0, dtype=torch.float64, **kw3)
    npo = len(POLES) if poles else 0
    s2 = yee2d.tm_state(NA, NB, torch.float64, n_poles=npo)
    s3 = yee3d.state3d(nx, ny, nz, torch.float64, n_poles=npo)
    Ez = mp["Ez"].clone()
    Ez[0], Ez[-1], Ez[:, 0], Ez[:, -1] = 0, 0, 0, 0     # PEC walls
    s2.update(Ez=Ez, Dz=mp["eps"] * Ez, Hx=mp["Hx"].clone(), Hy=mp["Hy"].clone())
    e, ha, hb = _names(v)
    for name, p2 in ((e, Ez), (ha, mp["Hx"]), (hb, mp["Hy"])):
        _put(v, s3[name], p2)
    for _ in range(N):
        s2 = yee2d.tm_step(s2, c2)
        s3 = yee3d.step3d(s3, c3)
    return s2, s3


def _rel(a, b):
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("v", [2, 0, 1], ids=["z", "x", "y"])
def test_invariant_mode_equals_2d_tm_with_media(v):
    mp = _maps(7 + v)
    s2, s3 = _run(v, mp)
    e, ha, hb = _names(v)
    errs = {"E": _rel(_plane(v, s3[e]), s2["Ez"]), "Ha": _rel(_plane(v, s3[ha]), s2["Hx"]),
            "Hb": _rel(_plane(v, s3[hb]), s2["Hy"]),
            "P": _rel(_plane(v, s3[f"P{e[1]}_cur"]), s2["P_cur"]),
            "Pp": _rel(_plane(v, s3[f"P{e[1]}_prev"]), s2["P_prev"])}
    assert all(x <= 1e-12 for x in errs.values()), errs
    # the other three components stay exactly zero, and the mode stays invariant along v
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        if f not in (e, ha, hb):
            assert not s3[f].any(), f
    assert torch.equal(s3[e].narrow(v, 0, 1), s3[e].narrow(v, 1, 1))


@pytest.mark.parametrize("sigma,poles", [(True, False), (False, True)], ids=["sigma", "poles"])
def test_each_medium_alone_and_that_it_matters(sigma, poles):
    """Conductivity alone and poles alone match the 2D scheme too, and each
    changes the fields by far more than the comparison tolerance."""
    mp = _maps(3)
    s2, s3 = _run(2, mp, sigma=sigma, poles=poles)
    assert _rel(_plane(2, s3["Ez"]), s2["Ez"]) <= 1e-12
    _, bare = _run(2, mp, sigma=False, poles=False)
    assert _rel(bare["Ez"], s3["Ez"]) > 1e-2


def test_uniform_conductor_decays_at_the_damping_rate():
    """With E and H uniform across a region far from the walls there is no
    curl there: each step multiplies E by d1 = (1 - s dt/2) / (1 + s dt/2)."""
    nx = ny = 12
    nz = 6
    s = 2.5
    c = yee3d.coeffs3d(nx, ny, nz, DX, torch.ones(nx, ny, nz), 0, dtype=torch.float64,
                       sigma_d=torch.full((nx, ny, nz), s, dtype=torch.float64))
    st = yee3d.state3d(nx, ny, nz, torch.float64)
    st["Ez"][1:-1, 1:-1, :] = 1.0
    dt = c["dt"]
    d1 = (1 - s * dt / 2) / (1 + s * dt / 2)
    for n in range(1, 4):
        st = yee3d.step3d(st, c)
        # the curl front moves one cell per step inward from the walls
        assert torch.allclose(st["Ez"][6, 6], torch.full((nz - 1,), d1 ** n, dtype=torch.float64),
                              rtol=1e-14, atol=0)
