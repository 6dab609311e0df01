"""3D boundary modes without a full CPML (meep_torch/core/stepper_3d.py:
174-215,328-361) on the HIP two-pass kernels, bit for bit against the
reference's own runs (tests/golden/absorber_3d_*.npz, which also pin the
oracle: tests/test_oracle_media_golden.py):

* abs: the matched-sigma absorber (n_abs 4, no CPML) with material sigma_d,
  sigma_b and a Lorentz + Drude pole pair: the absorber-path kernels
  (one division per curl, no psi, boundary sigma folded into the damping
  pairs);
* mixed: CPML on x and y with the absorber on z (identity profile there);
* pec: neither (absorber-path arithmetic in a PEC box).

Fields, pole stacks and CPML memory (mixed) are compared with torch.equal in
FP32 and FP64; the flux accumulator and spectrum within 1e-4 (FP32) and
1e-12 (FP64) relative, as the staged DFT flush sums in its own order."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, NZ, N = 22, 20, 18, 120
POLES2 = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]
CASES = {"abs": dict(npml=0, n_abs=4, media=True), "mixed": dict(npml=4, n_abs=4, axes=(2,), media=False),
         "pec": dict(npml=0, n_abs=0, media=False)}


def _t(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t if dtype is None else t.to(dtype)


def _run(case, tag, dtype, chunks=1):
    from meep_torch_amd.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    g = golden(f"absorber_3d_{tag}")
    o = CASES[case]
    dev = lambda k: _t(g[k]).cuda()
    kw = dict(sigma_d=dev("sigma_d"), sigma_b=dev("sigma_b"), poles=POLES2,
              pole_masks=[dev("m1"), dev("m2")]) if o["media"] else {}
    c = build_3d_coeffs(NX, NY, NZ, 0.05, eps=dev("eps"), npml=o["npml"], n_abs=o["n_abs"],
                        absorber_axes=o.get("axes", ()), dtype=dtype, device="cuda", **kw)
    assert (c.cpml is None) == (o["npml"] == 0)
    src = PointSource3D(c.dt, "Ez", 9, NY // 2, NZ // 2, GaussianPulse(1.0, 2.0, cutoff=2.0))
    assert np.array_equal(np.array([src.amplitude(t) for t in range(N)]), g["amp"])
    mon = FluxMonitor3D(c.dt, [0.8, 1.0, 1.2], normal="x", idx=14, ap=2, dtype=dtype)
    st = init_3d_state(NX, NY, NZ, npml=o["npml"], n_poles=2 if o["media"] else 0, dtype=dtype, device="cuda")
    print(f"[absorber3d] {case} {tag} chunks={chunks}", flush=True)
    if chunks == 1:
        st, (acc,) = run_3d(st, c, N, source=src, monitors=[mon])
        return g, st, acc, mon
    from meep_torch_amd.core.engine3d import run_fields
    from meep_torch_amd.core.stepper_3d import state_names
    work = {f: getattr(st, f).clone() for f in state_names(st)}
    raw, step = None, N // chunks
    for t0 in range(0, N, step):
        raw = run_fields(work, c, min(step, N - t0), src, [mon], step0=t0, reuse=raw, raw=True)
    from meep_torch_amd.core.engine2d import finish_outputs
    (acc,) = finish_outputs(raw, work)
    return g, st._replace(**work), acc, mon


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_absorber_modes_bitexact_vs_reference(case, tag, dtype):
    g, st, acc, mon = _run(case, tag, dtype)
    names = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
    if CASES[case]["media"]:
        names += ["Px_cur", "Px_prev", "Py_cur", "Pz_cur", "Pz_prev"]
    if case == "mixed":
        names += ["psi_Ez_y", "psi_Hy_z", "psi_Hx_y", "psi_Ey_x"]
    for f in names:
        got, ref = getattr(st, f).cpu(), _t(g[f"{case}_{f}"])
        assert torch.equal(got, ref), (case, f, float((got - ref).abs().max()))
    tol = 1e-4 if dtype == torch.float32 else 1e-12   # spectra: the staged flush sums in its own order
    ra, rf = _t(g[f"{case}_acc"]), _t(g[f"{case}_flux"])
    assert float((acc.cpu() - ra).norm() / ra.norm()) <= tol, case
    assert float((mon.flux(acc, 0.05).cpu() - rf).abs().max() / rf.abs().max()) <= tol, case
    assert float(st.Ez.abs().max()) > 0


def test_absorber_chunked_equals_one_call():
    """Four chained calls (the accumulators carried across them) reproduce
    the single call, and so the reference, bit for bit."""
    g, st, acc, _ = _run("abs", "f32", torch.float32, chunks=4)
    for f in ("Ex", "Hz", "Pz_cur"):
        assert torch.equal(getattr(st, f).cpu(), _t(g[f"abs_{f}"])), f
    ra = _t(g["abs_acc"])
    assert float((acc.cpu() - ra).norm() / ra.norm()) <= 1e-4
