"""TE anisotropic media (meep_torch/core/coeffs.py:390-400,425-455,
stepper_te.py:137-160) on the HIP two-pass kernels, against the reference's
own runs (golden tensor_te_*.npz, which also pin the oracle in
tests/test_oracle_media_golden.py):

* ten: an in-plane tensor eps (exx, exy, eyy) with an oriented pole
  (fn, gamma, (sxx, sxy, syy)) and a shared scalar pole: the D half stages
  D - sum P in scratch, a second pass forms E = eps^-1 (D - P) with the
  cross terms from four-point averages, and the oriented pole's P is also
  driven by the other component's E(n-1/2);
* tenk: the tensor eps under chi3 (the Jacobi passes start from the tensor E).

Fields, D and the pole stacks are compared with torch.equal in FP32 and
FP64; the flux accumulator within 1e-4 (FP32) / 1e-12 (FP64) relative, as
the staged DFT flush sums in its own order.  The hand-written adjoint
transposes the tensor product (and the chi2/chi3 passes after it) and the
oriented poles' cross terms."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, N = 48, 40, 240
POLES2 = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]


class _Src:
    def __init__(self, mask, amps):
        self.mask, self._a = mask, amps

    def amplitude(self, t):
        return float(self._a[t])


def _setup(case, tag, dtype):
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.stepper_te import init_te_state
    g = golden(f"tensor_te_{tag}")
    t = lambda k: torch.from_numpy(np.ascontiguousarray(g[k])).cuda()
    grid = Grid(nx=NX, ny=NY, dx=0.05, dtype=dtype, device="cuda")
    ten = (t("exx"), t("exy"), t("eyy"))
    if case == "ten":
        kw = dict(eps_tensor=ten, poles=POLES2[:1], pole_masks=[t("m2")],
                  poles_tensor=[(1.3, 0.2, (0.9, 0.4, 0.5))], pole_masks_tensor=[t("m1")])
    else:
        kw = dict(eps_tensor=ten, chi3=t("chi3"))
    m = build_te_coeffs(grid, eps_inf=t("eps"), npml=6, **kw)
    mon = FluxMonitorTE(grid, [0.6, 0.8, 1.0], ix=38, lo=7, hi=NY - 7)
    st = init_te_state(grid, n_poles=int(m.c1x.shape[0]), n_poles_y=int(m.c1y.shape[0]))
    return g, m, _Src(t(f"{case}_mask"), g[f"{case}_amp"]), mon, st


@pytest.mark.parametrize("case", ["ten", "tenk"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("chunk", [0, 70])
def test_tensor_eps_matches_reference(case, tag, dtype, chunk):
    from meep_torch_amd.core.run import run_te
    g, m, src, mon, st = _setup(case, tag, dtype)
    print(f"[tensor_te] {case} {tag} chunk={chunk}", flush=True)
    st, (acc,) = run_te(st, m, N, source=src, monitors=[mon], checkpoint_every=chunk)
    names = ["Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"]
    if case == "ten":
        names += ["Px_cur", "Px_prev", "Py_cur", "Py_prev"]
    for f in names:
        got, ref = getattr(st, f).cpu(), torch.from_numpy(g[f"{case}_{f}"])
        assert torch.equal(got, ref), (case, f, float((got - ref).abs().max()))
    ra = torch.from_numpy(g[f"{case}_acc"])
    tol = 1e-4 if dtype == torch.float32 else 1e-12
    assert float((acc.cpu() - ra).norm() / ra.norm()) <= tol


@pytest.mark.parametrize("kerr", [False, True, "oriented"], ids=["tensor", "tensor+chi3", "tensor+oriented"])
def test_tensor_eps_gradients_match_oracle_autograd(kerr):
    """The TE adjoint transposes the tensor-eps product E = ixx rhs_x +
    ixy_x avg(rhs_y) (its cross terms scattered back with the other
    average), after the chi2/chi3 passes when there are some: the gradients
    of a flux objective with respect to the tensor's exx, exy and eyy maps
    (through the inverse maps build_te_coeffs forms) equal CPU autograd
    through the oracle's steps (FP64)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    g = golden("tensor_te_f64")
    t = lambda k: torch.from_numpy(np.ascontiguousarray(g[k]))
    dt = torch.float64
    n = 160
    grid = Grid(nx=NX, ny=NY, dx=0.05, dtype=dt, device="cuda")
    ten = [t(k).cuda().requires_grad_(True) for k in ("exx", "exy", "eyy")]
    ori = kerr == "oriented"   # the "ten" case's oriented pole and scalar pole
    okw_cpu = (dict(poles=POLES2[:1], pole_masks=[t("m2")], poles_tensor=[(1.3, 0.2, (0.9, 0.4, 0.5))],
                    pole_masks_tensor=[t("m1")]) if ori else dict(chi3=t("chi3")) if kerr else {})
    kw = {k: ([x.cuda() for x in v] if isinstance(v, list) and k.startswith("pole_masks") else
              v.cuda() if torch.is_tensor(v) else v) for k, v in okw_cpu.items()}
    m = build_te_coeffs(grid, eps_inf=t("eps").cuda(), npml=6, eps_tensor=tuple(ten), **kw)
    mon = FluxMonitorTE(grid, [0.6, 0.8, 1.0], ix=38, lo=7, hi=NY - 7)
    case = "ten" if ori else "tenk"
    src = _Src(t(f"{case}_mask").cuda(), g[f"{case}_amp"])
    st0 = init_te_state(grid, n_poles=int(m.c1x.shape[0]), n_poles_y=int(m.c1y.shape[0]))
    _, (acc,) = run_te(st0, m, n, source=src, monitors=[mon], checkpoint_every=60)
    mon.flux(acc, 0.05).sum().backward()
    torch.set_num_threads(8)
    oten = [t(k).clone().requires_grad_(True) for k in ("exx", "exy", "eyy")]
    c = yee2d.te_coeffs(NX, NY, 0.05, t("eps"), 6, dtype=dt, eps_tensor=tuple(oten), **okw_cpu)
    ost = yee2d.te_state(NX, NY, dt, npx=int(m.c1x.shape[0]), npy=int(m.c1y.shape[0]))
    _, (oacc,) = yee2d.run2d("te", ost, c, n, amps=g[f"{case}_amp"][:n], mask=t(f"{case}_mask"),
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 38, "lo": 7, "hi": NY - 7,
                                        "freqs": [0.6, 0.8, 1.0]}])
    refs = torch.autograd.grad(yee2d.flux_from_acc("te_x", oacc, 0.05).sum(), oten)
    errs = [float((a.grad.cpu() - r).norm() / r.norm()) for a, r in zip(ten, refs)]
    print(f"\n[tensor grad kerr={kerr}] rel-L2 exx/exy/eyy {errs}")
    assert all(e <= 1e-9 for e in errs) and all(float(r.abs().max()) > 0 for r in refs)
