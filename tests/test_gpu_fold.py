"""Symmetry-folded 2D runs (meep_torch/core/coeffs.py:211-281,457-469,
stepper_tm.py:94-155, stepper_te.py:80-97) on the HIP two-pass kernels,
against the reference's own runs (golden fold_2d_*.npz, which also pin the
oracle: tests/test_oracle_fold_golden.py): TM cells cut by an even / odd y
mirror, an even / odd x mirror, both mirrors, Rotate2 and Rotate4 (each
phase), TE cells cut by even / odd y and x mirrors.  Every field, the
on-plane H rows Hx0 / Hy0 and the CPML memory are compared with torch.equal
(FP32 and FP64), the flux accumulator within 1e-4 (FP32) / 1e-12 (FP64)
relative, as the staged DFT flush sums in its own order; a chunked run
carries the on-plane rows across calls.  TE eps gradients through the cuts
match CPU autograd through the oracle, and so do TM gradients through even
mirror cuts; TM gradients through odd or Rotate cuts raise."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu
from test_oracle_fold_golden import FOLD_TE, FOLD_TM, FREQS, N, STEPS, te_npml, tm_npml  # noqa: F401

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


class _Src:
    def __init__(self, mask, amps):
        self.mask, self._a = mask, amps

    def amplitude(self, t):
        return float(self._a[t])


def _check(g, prefix, st, acc, names, dtype):
    for f in names:
        got, ref = getattr(st, f).cpu(), torch.from_numpy(g[f"{prefix}_{f}"])
        assert torch.equal(got, ref), (prefix, f, float((got - ref).abs().max()))
    ra = torch.from_numpy(g[f"{prefix}_acc"])
    tol = 1e-4 if dtype == torch.float32 else 1e-12
    assert float((acc.cpu() - ra).norm() / ra.norm()) <= tol, prefix


@pytest.mark.parametrize("name", sorted(FOLD_TM))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_folds_match_reference(name, tag, dtype):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = golden(f"fold_2d_{tag}")
    kw = FOLD_TM[name]
    grid = Grid(nx=N, ny=N, dx=0.05, dtype=dtype, device="cuda")
    m = build_tm_coeffs(grid, eps_inf=torch.from_numpy(g["eps"]).cuda(), npml=tm_npml(kw), **kw)
    mon = FluxMonitorTM(grid, FREQS, x_normal=True, idx=26, lo=0, hi=28)
    print(f"[fold] tm {name} {tag}", flush=True)
    src = _Src(torch.from_numpy(g[f"tm{name}_mask"]).cuda(), g["amp"])
    st, (acc,) = run_tm(init_tm_state(grid, **kw), m, STEPS, source=src, monitors=[mon],
                        checkpoint_every=70 if name in ("r4-", "xy") else 0)
    names = [k[len(f"tm{name}_"):] for k in g.keys() if k.startswith(f"tm{name}_") and not k.endswith(("_mask", "_acc"))]
    assert ("Hx0" in names) == (st.Hx0 is not None) and ("Hy0" in names) == (st.Hy0 is not None)
    _check(g, f"tm{name}", st, acc, names, dtype)


@pytest.mark.parametrize("name", sorted(FOLD_TE))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_te_folds_match_reference(name, tag, dtype):
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    g = golden(f"fold_2d_{tag}")
    kw = FOLD_TE[name]
    grid = Grid(nx=N, ny=N, dx=0.05, dtype=dtype, device="cuda")
    m = build_te_coeffs(grid, eps_inf=torch.from_numpy(g["eps"]).cuda(), npml=te_npml(kw), **kw)
    mon = FluxMonitorTE(grid, FREQS, ix=26, lo=0, hi=28)
    print(f"[fold] te {name} {tag}", flush=True)
    src = _Src(torch.from_numpy(g[f"te{name}_mask"]).cuda(), g["amp"])
    st, (acc,) = run_te(init_te_state(grid), m, STEPS, source=src, monitors=[mon])
    _check(g, f"te{name}", st, acc, ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"),
           dtype)


@pytest.mark.parametrize("name", sorted(FOLD_TE))
def test_te_fold_eps_gradients_match_oracle_autograd(name):
    """The TE adjoint transposes the even cut's on-plane row (Dx row 0 under
    y_mirror = +1, Dy column 0 under x_mirror = +1, both driven by
    2 Hz[0] / dx with no psi at the cut); odd cuts keep the zero row.  The
    eps gradient of the flux objective equals CPU autograd through the
    oracle's folded TE run (FP64)."""
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE
    from meep_torch_amd.core.run import run_te
    from meep_torch_amd.core.stepper_te import init_te_state
    from oracle import yee2d
    g = golden("fold_2d_f64")
    kw = FOLD_TE[name]
    grid = Grid(nx=N, ny=N, dx=0.05, dtype=torch.float64, device="cuda")
    eps = torch.from_numpy(g["eps"]).cuda().requires_grad_(True)
    m = build_te_coeffs(grid, eps_inf=eps, npml=te_npml(kw), **kw)
    mon = FluxMonitorTE(grid, FREQS, ix=26, lo=0, hi=28)
    print(f"[fold] te gradient {name}", flush=True)
    mask = torch.from_numpy(g[f"te{name}_mask"])
    _, (acc,) = run_te(init_te_state(grid), m, STEPS, source=_Src(mask.cuda(), g["amp"]), monitors=[mon],
                       checkpoint_every=70)
    mon.flux(acc, 0.05).sum().backward()
    torch.set_num_threads(8)
    e = torch.from_numpy(g["eps"]).clone().requires_grad_(True)
    c = yee2d.te_coeffs(N, N, 0.05, e, te_npml(kw), dtype=torch.float64, **kw)
    _, (oacc,) = yee2d.run2d("te", yee2d.te_state(N, N, torch.float64), c, STEPS, amps=g["amp"], mask=mask,
                             monitors=[{"type": "flux", "kind": "te_x", "idx": 26, "lo": 0, "hi": 28,
                                        "freqs": FREQS}])
    ref, = torch.autograd.grad(yee2d.flux_from_acc("te_x", oacc, 0.05).sum(), e)
    err = float((eps.grad.cpu() - ref).norm() / ref.norm())
    print(f"[fold] te gradient {name}: rel-L2 {err:.2e}", flush=True)
    assert err <= 1e-10, (name, err)


@pytest.mark.parametrize("name", ["ym+", "xm+"])
def test_tm_even_fold_eps_gradients_match_oracle_autograd(name):
    """Under an even TM cut the cut row / column of Dz is live with zero H
    beyond it and its psi arrays start at the cut; the adjoint's A phase
    follows the same live set and psi indexing.  The eps gradient of the
    flux objective equals CPU autograd through the oracle (FP64)."""
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    from oracle import yee2d
    g = golden("fold_2d_f64")
    kw = FOLD_TM[name]
    grid = Grid(nx=N, ny=N, dx=0.05, dtype=torch.float64, device="cuda")
    eps = torch.from_numpy(g["eps"]).cuda().requires_grad_(True)
    m = build_tm_coeffs(grid, eps_inf=eps, npml=tm_npml(kw), **kw)
    mon = FluxMonitorTM(grid, FREQS, x_normal=True, idx=26, lo=0, hi=28)
    print(f"[fold] tm gradient {name}", flush=True)
    mask = torch.from_numpy(g[f"tm{name}_mask"])
    _, (acc,) = run_tm(init_tm_state(grid, **kw), m, STEPS, source=_Src(mask.cuda(), g["amp"]), monitors=[mon],
                       checkpoint_every=70)
    mon.flux(acc, 0.05).sum().backward()
    torch.set_num_threads(8)
    e = torch.from_numpy(g["eps"]).clone().requires_grad_(True)
    c = yee2d.tm_coeffs(N, N, 0.05, e, tm_npml(kw), dtype=torch.float64, **kw)
    _, (oacc,) = yee2d.run2d("tm", yee2d.tm_state(N, N, torch.float64, **kw), c, STEPS, amps=g["amp"], mask=mask,
                             monitors=[{"type": "flux", "kind": "tm_x", "idx": 26, "lo": 0, "hi": 28,
                                        "freqs": FREQS}])
    ref, = torch.autograd.grad(yee2d.flux_from_acc("tm_x", oacc, 0.05).sum(), e)
    err = float((eps.grad.cpu() - ref).norm() / ref.norm())
    print(f"[fold] tm gradient {name}: rel-L2 {err:.2e}", flush=True)
    assert err <= 1e-10, (name, err)


def test_folds_refuse_gradients_and_batches():
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.stepper_tm import init_tm_state
    grid = Grid(nx=N, ny=N, dx=0.05, dtype=torch.float64, device="cuda")
    eps = torch.ones(N, N, dtype=torch.float64, device="cuda", requires_grad=True)
    m = build_tm_coeffs(grid, eps_inf=eps, npml=((6, 6), (0, 6)), y_mirror=-1)
    src = _Src(torch.zeros(N, N, dtype=torch.float64, device="cuda"), np.zeros(10))
    with pytest.raises(NotImplementedError, match="symmetry"):
        run_tm(init_tm_state(grid, y_mirror=-1), m, 10, source=src, monitors=[], checkpoint_every=5)
    with pytest.raises(NotImplementedError, match="symmetry"):
        BatchedTM([(m._replace(eps_inf=eps.detach()), None, [])], 10, checkpoint_every=0, adjoint=False)
