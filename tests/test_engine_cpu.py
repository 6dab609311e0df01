"""CPU-side checks of the engine boundary: the C-ABI library loads and exports
every entry point the headers under include/ declare; host logic (profiles, shell
bounds, source boxes, feature-flag errors) behaves like the reference."""
import ctypes
import os
import re

import pytest
import torch

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("meep_fdtd.h", "meep_surrogate.h")]


def _declared():
    text = "\n".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w ]+?\*?\s*\b(mt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from meep_torch_amd import _engine
    L = _engine.lib()
    names = _declared()
    assert {"mt_tm2d_run", "mt_te2d_run", "mt_yee3d_run", "mt_last_error", "mt_gemm_bf16"} <= set(names)
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == sorted(os.path.basename(h) for h in HEADERS)
    for n in names:
        assert hasattr(L, n), f"libmeepfdtd.so does not export {n}"
    assert L.mt_abi_version() == _engine.ABI_VERSION


def test_ctypes_structs_match_the_compiled_layout():
    """Every ABI struct the Python binding declares has the size (and the
    late fields the offsets) the library was compiled with: a mismatch would
    hand the kernels wrong pointers."""
    from meep_torch_amd import _engine as E
    from meep_torch_amd.batch import BatchItem
    L = E.lib()
    buf = (ctypes.c_longlong * 64)()
    n = L.mt_abi_layout(buf, 64)
    assert n == 26
    sizes = [E.PmlAxis, E.Source2D, E.TM2D, E.TE2D, E.Yee3D, E.Source3D, E.DftTerm, E.Monitor, E.TM2DAdj,
             E.TE2DAdj, E.Yee3DAdj, BatchItem]
    for q, st in enumerate(sizes):
        assert buf[q] == ctypes.sizeof(st), (st.__name__, buf[q], ctypes.sizeof(st))
    offs = [E.TM2D.hsrc.offset, E.TM2D.hcoef.offset, E.TE2D.tape.offset, E.TE2D.hsrc.offset,
            E.TE2D.hcoef.offset, E.Yee3D.alt.offset, E.Monitor.acc.offset, E.TM2D.n_poles.offset,
            E.TM2D.h2y.offset, E.TE2D.npx.offset, E.TE2D.hz2.offset, E.TE2D.ky.offset,
            E.Yee3D.damp.offset, E.Yee3D.P_prev.offset]
    assert list(buf[12:26]) == offs


def test_error_path_without_gpu_work():
    """Invalid descriptors are rejected on the host before any launch."""
    from meep_torch_amd import _engine as E
    L = E.lib()
    d = E.TM2D()
    d.nx, d.ny = 2, 2
    rc = L.mt_tm2d_run(ctypes.byref(d), None, None, None, 0, None, 0, 0, 0, 1, None)
    assert rc != 0 and b"3x3" in L.mt_last_error()
    rc = L.mt_tm2d_run(None, None, None, None, 0, None, 0, 0, 0, 1, None)
    assert rc != 0


def test_shell_bounds_match_profile():
    from meep_torch_amd._engine import shell_bounds
    from meep_torch_amd.core.pml import cpml_profiles
    b, a, _ = cpml_profiles(255, 16, 0.5, 0.025, 0.05, dtype=torch.float32)
    lo, hi = shell_bounds(b, a)
    assert torch.all(b[lo:hi] == 1) and torch.all(a[lo:hi] == 0)
    assert b[lo - 1] != 1 and b[hi] != 1
    b, a, _ = cpml_profiles(50, 0, 0.0, 0.025, 0.05)
    assert shell_bounds(b, a) == (0, 50)


def test_package_profiles_bitexact_vs_golden():
    import json
    from conftest import golden
    from meep_torch_amd.core.pml import cpml_profiles
    g = golden("pml")
    for ci, (n, npml, off) in enumerate(json.loads(bytes(g["combos_json"]).decode())):
        npml = tuple(npml) if isinstance(npml, list) else npml
        for tag, dt in (("f32", torch.float32), ("f64", torch.float64)):
            b, a, k = cpml_profiles(n, npml, off, 0.025, 0.05, dtype=dt)
            assert torch.equal(b, torch.from_numpy(g[f"{tag}_{ci}_b"]))
            assert torch.equal(a, torch.from_numpy(g[f"{tag}_{ci}_a"]))


def test_mask_box():
    from meep_torch_amd._engine import mask_box
    m = torch.zeros(10, 8)
    m[3, 2:6] = 2.0
    (i0, j0), (i1, j1), w = mask_box(m)
    assert (i0, i1, j0, j1) == (3, 4, 2, 6) and torch.equal(w, m[3:4, 2:6])
    assert mask_box(torch.zeros(4, 4)) is None


@pytest.mark.parametrize("te,kw,flag", [(False, {"y_mirror": -1, "mu": 2.0}, "odd mirror"),
                                        (False, {"y_rot2": 1, "sigma_b": 0.5}, "Rotate2"),
                                        (False, {"m_poles": [(1.0, 0.1, 1.0)], "mu": 2.0}, "m_poles")])
def test_unsupported_flags_raise(te, kw, flag):
    """The combinations the reference refuses raise by name: an odd mirror or
    Rotate cut with sigma_b / mu (no coefficients on the on-plane H row,
    coeffs.py:239-243) and magnetic poles with sigma_b / mu
    (coeffs.py:178-179)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    g = Grid(nx=16, ny=12, dx=0.05, dtype=torch.float32, device="cpu")
    with pytest.raises(NotImplementedError, match=flag):
        (build_te_coeffs if te else build_tm_coeffs)(g, npml=2, **kw)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_te_tensor_coefficients_equal_the_oracle(dtype):
    """build_te_coeffs's tensor-eps inverse components and oriented-pole
    stacks (c1..c3 with the appended poles, c3xy / c3yx) are the oracle's
    bit for bit."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from oracle import yee2d
    gm = torch.Generator().manual_seed(9)
    r = lambda: torch.rand(16, 12, generator=gm, dtype=torch.float64).to(dtype)
    ten = (2.0 + r(), 0.3 * r(), 1.5 + r())
    kw = dict(eps_tensor=ten, poles=[(1.1, 0.15, 1.3)], pole_masks=[r()],
              poles_tensor=[(1.3, 0.2, (0.9, 0.4, 0.5)), (0.8, 0.1, (0.2, -0.3, 0.6), 0.7)],
              pole_masks_tensor=[r(), r()])
    g = Grid(nx=16, ny=12, dx=0.05, dtype=dtype, device="cpu")
    c = build_te_coeffs(g, npml=2, **kw)
    oc = yee2d.te_coeffs(16, 12, 0.05, torch.ones(16, 12), 2, dtype=dtype, **kw)
    for name in ("inv_xx_x", "inv_xy_x", "inv_yy_y", "inv_xy_y", "c1x", "c2x", "c3x", "c1y", "c2y", "c3y",
                 "c3xy", "c3yx"):
        assert torch.equal(getattr(c, name), getattr(oc, name)), name
    assert tuple(c.c3xy.shape) == (3, 15, 12) and not c.c3xy[0].any()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_tm_magnetic_pole_coefficients_equal_the_oracle(dtype):
    """build_tm_coeffs's magnetic pole arrays (mc1..mc3 on the Hx / Hy grids)
    are the oracle's bit for bit (the oracle's are pinned to the reference's
    runs by tests/test_oracle_media_golden.py)."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from oracle import yee2d
    gm = torch.Generator().manual_seed(5)
    masks = [torch.rand(16, 12, generator=gm, dtype=torch.float64).to(dtype) for _ in range(2)]
    poles = [(1.0, 0.2, 1.5), (0.7, 0.1, 0.8)]
    g = Grid(nx=16, ny=12, dx=0.05, dtype=dtype, device="cpu")
    c = build_tm_coeffs(g, npml=2, m_poles=poles, m_pole_masks=masks)
    oc = yee2d.tm_coeffs(16, 12, 0.05, torch.ones(16, 12), 2, dtype=dtype, m_poles=poles, m_pole_masks=masks)
    for name in ("mc1x", "mc2x", "mc3x", "mc1y", "mc2y", "mc3y"):
        assert torch.equal(getattr(c, name), getattr(oc, name)), name
    assert tuple(c.mc3x.shape) == (2, 16, 11) and tuple(c.mc3y.shape) == (2, 15, 12)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_tm_media_coefficients_equal_the_oracle(dtype):
    """build_tm_coeffs's pole, chi2/chi3 and sigma/mu arrays are the oracle's,
    bit for bit (the oracle's are pinned by tests/test_oracle_media_cpu.py)."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from oracle import yee2d
    nx, ny = 24, 18
    gen = torch.Generator().manual_seed(1)
    eps = 1.0 + torch.rand(nx, ny, generator=gen, dtype=torch.float64)
    m1 = (torch.rand(nx, ny, generator=gen, dtype=torch.float64) > 0.5).double()
    kw = dict(poles=[(1.2, 0.1, 1.5), (0.0, 0.3, 2.0, 0.9)], pole_masks=[m1, 1.0 - m1],
              chi3=0.02 * m1, chi2=0.01 * (1 - m1), sigma_d=0.5 * m1, sigma_b=0.2 * m1, mu=1.0 + m1)
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=dtype, device="cpu")
    m = build_tm_coeffs(g, eps_inf=eps, npml=3, **{k: (v.to(dtype) if torch.is_tensor(v) else
                                                       [x.to(dtype) for x in v] if k == "pole_masks" else v)
                                                    for k, v in kw.items()})
    o = yee2d.tm_coeffs(nx, ny, 0.05, eps, 3, dtype=dtype, **{k: (v.to(dtype) if torch.is_tensor(v) else
                                                                  [x.to(dtype) for x in v] if k == "pole_masks"
                                                                  else v) for k, v in kw.items()})
    assert m.has_kerr and m.has_chi2 and o.kerr
    for name in ("c1", "c2", "c3", "chi2", "chi3", "d1", "d2", "h1x", "h2x", "h1y", "h2y"):
        assert torch.equal(getattr(m, name), getattr(o, name)), name


def test_run_requires_hip_device():
    """No CPU fallback: CPU tensors are refused loudly."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=16, ny=12, dx=0.05, dtype=torch.float32, device="cpu")
    with pytest.raises(RuntimeError, match="HIP engine"):
        run_tm(init_tm_state(g), build_tm_coeffs(g, npml=2), 3)


def test_grid_validation():
    from meep_torch_amd.core.grid import Grid
    with pytest.raises(ValueError):
        Grid(nx=2, ny=8, dx=0.1)
    with pytest.raises(ValueError):
        Grid(nx=8, ny=8, dx=0.1, courant=0.8)
    assert Grid(nx=8, ny=8, dx=0.1, courant=0.5, device="cpu").dt == 0.05


@pytest.mark.parametrize("what", ["sigma_d", "sigma_b", "pole_masks"])
def test_3d_adjoint_refuses_material_gradients(what):
    """run_3d_adjoint differentiates with respect to eps only: a sigma_d,
    sigma_b or pole-mask tensor that requires grad raises by name (as the 2D
    adjoint does for non-eps coefficients) before any device work."""
    from meep_torch_amd.core.adjoint3d import run_3d_adjoint
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    n = 12
    m = torch.zeros(n, n, n, dtype=torch.float64)
    m[4:8, 4:8, 4:8] = 1.0
    kw = {}
    if what == "pole_masks":
        kw = dict(poles=[(1.0, 0.1, 1.0)], pole_masks=[m.clone().requires_grad_(True)])
    else:
        kw = {what: m.clone().requires_grad_(True)}
    c = build_3d_coeffs(n, n, n, 0.05, eps=torch.ones(n, n, n, dtype=torch.float64), npml=2,
                        dtype=torch.float64, device="cpu", **kw)
    st = init_3d_state(n, n, n, npml=2, n_poles=1 if what == "pole_masks" else 0, dtype=torch.float64,
                       device="cpu")
    with pytest.raises(NotImplementedError, match="sigma_d|sigma_b|pole"):
        run_3d_adjoint(c, 4, 2, None, [], lambda accs: 0.0, st)
