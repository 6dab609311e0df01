"""Pin the oracle's media terms to the reference bit for bit: golden vectors
written by the reference itself (tests/golden/make_golden.py, cases media_tm,
media_te, media_3d) must be reproduced exactly by the oracle's restatement,
in FP32 and FP64.

* TM and TE 48 x 40 cells with a Lorentz and a Drude pole (TE: one shared
  pole, one on Ex only, one on Ey only), chi2 and chi3 regions, sigma_d,
  sigma_b and mu, under a line source ("all"), and with chi2 + chi3 alone
  under a strong drive ("nl"), where the three-pass fixed point
  (stepper_tm.py:228-231, stepper_te.py:161-175) changes the fields by far
  more than rounding;
* TM magnetic Lorentz poles, with electric poles and chi3 and under a TFSF
  magnetic current (mag_tm);
* TE tensor eps with an oriented pole, and under chi3 (tensor_te);
* 3D 22 x 20 x 18 with sigma_d, sigma_b and two poles under a point source,
  and the boundary modes without a full CPML (absorber_3d).

Bit-exact equality pins every coefficient, sign and operation order of the
nonlinear and dispersive updates (a wrong chi2 / chi3 factor, a dropped pole
term or a swapped average fails here); the engine's kernels equal the oracle
bit for bit on the GPU (tests/test_gpu_media*.py), so these fixtures pin
them too."""
import numpy as np
import pytest
import torch

from conftest import golden
from oracle import yee2d, yee3d

torch.set_num_threads(8)
POLES2 = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]
FREQS = [0.6, 0.8, 1.0]
NX, NY, N = 48, 40, 240


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def _eq(a, b, what):
    b = _t(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    assert torch.equal(a, b), (what, float((a - b).abs().max()))


def _kw(g, case, te):
    mp = {k: _t(g[k]) for k in ("m1", "m2", "chi2", "chi3", "sigma_d", "sigma_b", "mu")}
    if case == "nl":
        return dict(chi2=mp["chi2"], chi3=mp["chi3"])
    kw = dict(chi2=mp["chi2"], chi3=mp["chi3"], sigma_d=mp["sigma_d"], sigma_b=mp["sigma_b"], mu=mp["mu"])
    if te:
        kw.update(poles=POLES2[:1], pole_masks=[mp["m1"]], poles_x=POLES2[1:], pole_masks_x=[mp["m2"]],
                  poles_y=[(1.6, 0.2, 0.7)], pole_masks_y=[mp["m2"]])
    else:
        kw.update(poles=POLES2, pole_masks=[mp["m1"], mp["m2"]])
    return kw


@pytest.mark.parametrize("case", ["all", "nl"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_media_bitexact(tag, dtype, case):
    g = golden(f"media_tm_{tag}")
    kw = _kw(g, case, te=False)
    c = yee2d.tm_coeffs(NX, NY, 0.05, _t(g["eps"]), 6, dtype=dtype, **kw)
    n_poles = len(kw.get("poles", ()))
    mons = [{"type": "flux", "kind": "tm_x", "idx": 38, "lo": 7, "hi": NY - 7, "freqs": FREQS}]
    s, acc = yee2d.run2d("tm", yee2d.tm_state(NX, NY, dtype, n_poles=n_poles), c, N, amps=g[f"{case}_amp"],
                         mask=_t(g[f"{case}_mask"]), monitors=mons)
    names = ["Ez", "Dz", "Hx", "Hy", "psi_Ezy", "psi_Ezx", "psi_Hyx", "psi_Hxy"]
    if n_poles:
        names += ["P_cur", "P_prev"]
    for f in names:
        _eq(s[f], g[f"{case}_{f}"], f)
    _eq(acc[0], g[f"{case}_acc"], "acc")


@pytest.mark.parametrize("case", ["all", "nl"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_te_media_bitexact(tag, dtype, case):
    g = golden(f"media_te_{tag}")
    kw = _kw(g, case, te=True)
    c = yee2d.te_coeffs(NX, NY, 0.05, _t(g["eps"]), 6, dtype=dtype, **kw)
    npx, npy = (2, 2) if case == "all" else (0, 0)
    mons = [{"type": "flux", "kind": "te_x", "idx": 38, "lo": 7, "hi": NY - 7, "freqs": FREQS}]
    s, acc = yee2d.run2d("te", yee2d.te_state(NX, NY, dtype, npx=npx, npy=npy), c, N, amps=g[f"{case}_amp"],
                         mask=_t(g[f"{case}_mask"]), monitors=mons)
    names = ["Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"]
    if npx:
        names += ["Px_cur", "Px_prev", "Py_cur", "Py_prev"]
    for f in names:
        _eq(s[f], g[f"{case}_{f}"], f)
    _eq(acc[0], g[f"{case}_acc"], "acc")


@pytest.mark.parametrize("case", ["ten", "tenk"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_te_tensor_eps_and_oriented_poles_bitexact(tag, dtype, case):
    """TE anisotropy (coeffs.py:390-400,425-455, stepper_te.py:137-160):
    E = eps^-1 (D - sum P) with the off-diagonal inverse applied to the
    four-point average of the other component's rhs, an oriented pole whose
    P is also driven by the other component's E(n-1/2), and the tensor eps
    under the chi3 passes (which start from the tensor E), reproduced bit
    for bit from the reference's runs.  A transposed inverse, a dropped
    cross term or the cross drive from E(n+1/2) fails here."""
    g = golden(f"tensor_te_{tag}")
    ten = (_t(g["exx"]), _t(g["exy"]), _t(g["eyy"]))
    if case == "ten":
        kw = dict(eps_tensor=ten, poles=POLES2[:1], pole_masks=[_t(g["m2"])],
                  poles_tensor=[(1.3, 0.2, (0.9, 0.4, 0.5))], pole_masks_tensor=[_t(g["m1"])])
    else:
        kw = dict(eps_tensor=ten, chi3=_t(g["chi3"]))
    c = yee2d.te_coeffs(NX, NY, 0.05, _t(g["eps"]), 6, dtype=dtype, **kw)
    npx = 2 if case == "ten" else 0
    mons = [{"type": "flux", "kind": "te_x", "idx": 38, "lo": 7, "hi": NY - 7, "freqs": FREQS}]
    s, acc = yee2d.run2d("te", yee2d.te_state(NX, NY, dtype, npx=npx, npy=npx), c, N, amps=g[f"{case}_amp"],
                         mask=_t(g[f"{case}_mask"]), monitors=mons)
    names = ["Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Hzx"] + (["Px_cur", "Py_prev"] if npx else [])
    for f in names:
        _eq(s[f], g[f"{case}_{f}"], f"{case} {f}")
    _eq(acc[0], g[f"{case}_acc"], "acc")


def test_nonlinear_terms_matter():
    """The strong-drive fixture is a real pin of chi2 / chi3: without them
    the reference's own fields differ from the fixture by far more than
    rounding (FP64 oracle run with the nonlinear maps zeroed)."""
    g = golden("media_tm_f64")
    c = yee2d.tm_coeffs(NX, NY, 0.05, _t(g["eps"]), 6, dtype=torch.float64)
    s, _ = yee2d.run2d("tm", yee2d.tm_state(NX, NY, torch.float64), c, N, amps=g["nl_amp"],
                       mask=_t(g["nl_mask"]), monitors=[])
    ref = _t(g["nl_Ez"])
    assert float((s["Ez"] - ref).norm() / ref.norm()) > 1e-3


@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_3d_media_bitexact(tag, dtype):
    g = golden(f"media_3d_{tag}")
    nx, ny, nz, npml, n = 22, 20, 18, 4, 120
    c = yee3d.coeffs3d(nx, ny, nz, 0.05, _t(g["eps"]), npml, dtype, sigma_d=_t(g["sigma_d"]),
                       sigma_b=_t(g["sigma_b"]), poles=POLES2, pole_masks=[_t(g["m1"]), _t(g["m2"])])
    mons = [{"type": "flux", "normal": "x", "idx": 15, "freqs": [0.8, 1.0, 1.2], "ap": npml + 1}]
    s, acc = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype, n_poles=2), c, n, list(g["amp"]),
                         ("Ez", 7, ny // 2, nz // 2), mons)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "psi_Ez_x", "psi_Hy_z", "psi_Hx_y", "Px_cur", "Py_prev",
              "Pz_cur"):
        _eq(s[f], g[f], f)
    _eq(acc[0], g["acc"], "acc")
    _eq(yee3d.flux_plane(acc[0], 0.05), g["flux"], "flux")


def test_3d_lean_poles_bitexact():
    """The lean in-place stepper with two ADE poles on masks (stepper_3d_lean.py:
    143-165,207-231), reproduced bit for bit from the reference's run: the pole
    advances from E(n) before the E update and E = E + ((curl dt + sum P(n))
    - sum P(n+1)) inv_eps, the recurrence taking the reference's in-place
    mul_ / addcmul_ (each multiply-add rounded once).  The functional pole
    form (d1 (E eps + sum P) + d2 curl - sum P') / eps, the lean form without
    the pole sums, or an unfused recurrence fails here."""
    g = golden("lean_poles_3d_f32")
    nx, ny, nz, npml, n = 22, 20, 18, 4, 120
    dtype = torch.float32
    c = yee3d.coeffs3d(nx, ny, nz, 0.05, _t(g["eps"]), npml, dtype, poles=POLES2,
                       pole_masks=[_t(g["m1"]), _t(g["m2"])])
    mons = [{"type": "flux", "normal": "x", "idx": 15, "freqs": [0.8, 1.0, 1.2], "ap": npml + 1}]
    s, acc = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype, n_poles=2), c, n, list(g["amp"]),
                         ("Ez", 7, ny // 2, nz // 2), mons, lean=True)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        _eq(s[f], g[f], f)
    for a in "xyz":
        _eq(s[f"P{a}_cur"], g[f"P{a}_new"], f"P{a} new")
        _eq(s[f"P{a}_prev"], g[f"P{a}_old"], f"P{a} old")
    _eq(acc[0], g["acc"], "acc")
    # the functional pole form differs from the lean one on the same run
    s2, _ = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype, n_poles=2), c, 40, list(g["amp"])[:40],
                        ("Ez", 7, ny // 2, nz // 2), [], lean=False)
    s3, _ = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype, n_poles=2), c, 40, list(g["amp"])[:40],
                        ("Ez", 7, ny // 2, nz // 2), [], lean=True)
    assert not torch.equal(s2["Ez"], s3["Ez"])


MPOLES2 = [(1.0, 0.2, 1.5), (0.7, 0.1, 0.8)]
MAG_FIELDS = ("Ez", "Dz", "Hx", "Hy", "Bx", "By", "Mx_cur", "Mx_prev", "My_cur", "My_prev", "psi_Ezy", "psi_Hyx")


@pytest.mark.parametrize("case", ["mag", "magh"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_magnetic_poles_bitexact(case, tag, dtype):
    """Magnetic Lorentz poles in 2D TM (coeffs.py:177-196, stepper_tm.py:
    103-106,180-193,209-224), reproduced bit for bit from the reference's
    runs: B integrates the curl, the TFSF Hy current goes into B, and
    H = B - sum M' with M' = mc1 M + mc2 M_prev + mc3 H(n-1/2); with
    electric poles and chi3 in the same cell (mag) and alone under a
    magnetic current (magh).  A dropped mc3 term, H(n+1/2) in place of
    H(n-1/2) in the recurrence, or the current into H instead of B fails
    here."""
    g = golden(f"mag_tm_{tag}")
    nx, ny, n = 48, 40, 240
    kw = dict(m_poles=MPOLES2, m_pole_masks=[_t(g["mm1"]), _t(g["mm2"])])
    if case == "mag":
        kw.update(poles=POLES2, pole_masks=[_t(g["m1"]), _t(g["m2"])], chi3=_t(g["chi3"]))
    c = yee2d.tm_coeffs(nx, ny, 0.05, _t(g["eps"]), 6, dtype=dtype, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 38, "lo": 7, "hi": ny - 7, "freqs": FREQS}]
    st = yee2d.tm_state(nx, ny, dtype, n_poles=2 if case == "mag" else 0, n_mpoles=2)
    hk = {} if case == "mag" else dict(hamps=g["magh_h_amp"], hmasks={"hy_mask": _t(g["magh_hy_mask"])})
    s, acc = yee2d.run2d("tm", st, c, n, amps=g["mag_amp"], mask=_t(g["mag_mask"]), monitors=mons, **hk)
    for f in MAG_FIELDS + (("P_cur",) if case == "mag" else ()):
        _eq(s[f], g[f"{case}_{f}"], f"{case} {f}")
    _eq(acc[0], g[f"{case}_acc"], f"{case} acc")


ABS_CASES = {"abs": dict(npml=0, n_abs=4, media=True), "mixed": dict(npml=4, n_abs=4, axes=(2,), media=False),
             "pec": dict(npml=0, n_abs=0, media=False)}


@pytest.mark.parametrize("case", sorted(ABS_CASES))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_3d_absorber_modes_bitexact(case, tag, dtype):
    """Boundary modes without a full CPML (stepper_3d.py:174-215,328-361),
    reproduced bit for bit from the reference's own run: the matched-sigma
    absorber (one division per curl, sig_bound added to the material sigma,
    poles), the mixed mode (identity profile on the absorber axis) and the
    absorber-path arithmetic with no boundary sigma at all."""
    g = golden(f"absorber_3d_{tag}")
    o = ABS_CASES[case]
    nx, ny, nz, n = 22, 20, 18, 120
    kw = dict(sigma_d=_t(g["sigma_d"]), sigma_b=_t(g["sigma_b"]), poles=POLES2,
              pole_masks=[_t(g["m1"]), _t(g["m2"])]) if o["media"] else {}
    c = yee3d.coeffs3d(nx, ny, nz, 0.05, _t(g["eps"]), o["npml"], dtype, n_abs=o["n_abs"],
                       absorber_axes=o.get("axes", ()), absorber=(o["npml"] == 0), **kw)
    assert c["absorber"] == (o["npml"] == 0)
    mons = [{"type": "flux", "normal": "x", "idx": 14, "freqs": [0.8, 1.0, 1.2], "ap": 2}]
    s, acc = yee3d.run3d(yee3d.state3d(nx, ny, nz, dtype, n_poles=2 if o["media"] else 0), c, n,
                         list(g["amp"]), ("Ez", 9, ny // 2, nz // 2), mons)
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz") + (("Px_cur", "Pz_prev") if o["media"] else ()):
        _eq(s[f], g[f"{case}_{f}"], f"{case} {f}")
    _eq(acc[0], g[f"{case}_acc"], f"{case} acc")
    _eq(yee3d.flux_plane(acc[0], 0.05), g[f"{case}_flux"], f"{case} flux")


def test_3d_absorber_is_not_the_cpml_arithmetic():
    """The absorber path's single division per curl is a different rounding
    from the CPML path's per-derivative division: the same PEC box run with
    CPML arithmetic (npml 0: b = 1, a = 0, psi zero) differs from the
    reference's absorber-path fixture in FP32, so the bit-exact test above
    sees the arithmetic, not only the physics."""
    g = golden("absorber_3d_f32")
    nx, ny, nz, n = 22, 20, 18, 120
    c = yee3d.coeffs3d(nx, ny, nz, 0.05, _t(g["eps"]), 0, torch.float32)
    assert not c["absorber"]
    s, _ = yee3d.run3d(yee3d.state3d(nx, ny, nz, torch.float32), c, n, list(g["amp"]),
                       ("Ez", 9, ny // 2, nz // 2), [])
    ref = _t(g["pec_Ez"])
    assert not torch.equal(s["Ez"], ref)
    assert float((s["Ez"] - ref).norm() / ref.norm()) < 1e-4   # same physics, other rounding


def test_tfsf_mode_source_run_bitexact():
    """A ModeSourceTM (TFSF eigenmode source) run of the reference, replayed
    by the oracle from the reference's own masks and electric / magnetic
    amplitude sequences: fields and both flux spectra bit for bit (FP64),
    pinning the B-phase magnetic current (stepper_tm.py:169-193) and the
    h_masks / h_amplitude protocol of run.py:17-23."""
    g = golden("tfsf_tm_f64")
    nx, ny, n = 160, 64, 600
    c = yee2d.tm_coeffs(nx, ny, 0.05, _t(g["eps"]), 12, dtype=torch.float64)
    mons = [{"type": "flux", "kind": "tm_x", "idx": i, "lo": 14, "hi": ny - 14, "freqs": [0.5, 0.6, 0.7]}
            for i in (120, 20)]
    s, acc = yee2d.run2d("tm", yee2d.tm_state(nx, ny, torch.float64), c, n, amps=g["amp"], mask=_t(g["mask"]),
                         monitors=mons, hamps=g["hamp"], hmasks={"hy_mask": _t(g["src_hy_mask"])})
    for f in ("Ez", "Dz", "Hx", "Hy"):
        _eq(s[f], g[f], f)
    _eq(acc[0], g["acc0"], "acc0")
    _eq(acc[1], g["acc1"], "acc1")


def test_package_mode_source_matches_the_reference_sheets():
    """The package's ModeSourceTM builds the reference's sheets: the same
    neff, electric / magnetic masks and amplitude sequences within the mode
    solver's rounding (the profiles come from an eigen-solve of its own)."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.ports.mode_source import ModeSourceTM
    from meep_torch_amd.ports.port import Port
    gd = golden("tfsf_tm_f64")
    nx, ny, n = 160, 64, 600
    g = Grid(nx=nx, ny=ny, dx=0.05, dtype=torch.float64, device="cpu")
    m = build_tm_coeffs(g, eps_inf=_t(gd["eps"]), npml=12)
    src = ModeSourceTM(g, m, Port(axis="x", index=40, lo=14, hi=ny - 14, direction=+1),
                       GaussianPulse(0.6, 0.3), 0.6)
    assert abs(src.neff - float(gd["neff"])) <= 1e-10
    for mine, ref in ((src.mask, gd["mask"]), (src.h_masks["src_hy_mask"], gd["src_hy_mask"])):
        ref = _t(ref)
        # the eigenvector's overall sign is fixed by the same convention
        assert float((mine.cpu() - ref).norm() / ref.norm()) <= 1e-9
    amps = np.array([src.amplitude(t) for t in range(n)])
    hamps = np.array([src.h_amplitude(t) for t in range(n)])
    assert np.array_equal(amps, gd["amp"])
    assert np.abs(hamps - gd["hamp"]).max() <= 1e-12 * np.abs(gd["hamp"]).max()
