"""TM magnetic Lorentz poles (meep_torch/core/coeffs.py:177-196,
stepper_tm.py:103-106,180-193,209-224) on the HIP two-pass kernels, against
the reference's own runs (golden mag_tm_*.npz, which also pin the oracle in
tests/test_oracle_media_golden.py):

* mag: two magnetic poles with cell masks, two electric poles and chi3 in
  one cell under a line source;
* magh: the magnetic poles alone under a line source and a TFSF-style Hy
  current, which the B phase injects into B.

Fields, B, the M and P stacks and CPML memory are compared with torch.equal
(FP32 and FP64); the flux accumulator within 1e-4 (FP32) / 1e-12 (FP64)
relative, as the staged DFT flush sums in its own order.  Also: chunked runs
carry B and the M stacks across calls, and the eps gradient through the
magnetic poles (the hand-written TM adjoint transposes the B / M recurrence)
equals CPU autograd through the oracle."""
import numpy as np
import pytest
import torch

from conftest import golden, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

NX, NY, N = 48, 40, 240
POLES2 = [(1.1, 0.15, 1.3), (0.0, 0.4, 2.0, 0.9)]
MPOLES2 = [(1.0, 0.2, 1.5), (0.7, 0.1, 0.8)]
NAMES = ("Ez", "Dz", "Hx", "Hy", "Bx", "By", "Mx_cur", "Mx_prev", "My_cur", "My_prev", "psi_Ezy", "psi_Ezx",
         "psi_Hyx", "psi_Hxy")


class _HSource:
    def __init__(self, mask, amps, hy_mask, h_amps):
        self.mask, self._a, self.h_masks, self._h = mask, amps, {"src_hy_mask": hy_mask}, h_amps

    def amplitude(self, t):
        return float(self._a[t])

    def h_amplitude(self, t):
        return float(self._h[t])


def _setup(case, tag, dtype):
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = golden(f"mag_tm_{tag}")
    t = lambda k: torch.from_numpy(np.ascontiguousarray(g[k])).cuda()
    grid = Grid(nx=NX, ny=NY, dx=0.05, dtype=dtype, device="cuda")
    kw = dict(m_poles=MPOLES2, m_pole_masks=[t("mm1"), t("mm2")])
    if case == "mag":
        kw.update(poles=POLES2, pole_masks=[t("m1"), t("m2")], chi3=t("chi3"))
    m = build_tm_coeffs(grid, eps_inf=t("eps"), npml=6, **kw)
    src = _HSource(t("mag_mask"), g["mag_amp"], t("magh_hy_mask"), g["magh_h_amp"])
    if case == "mag":
        src.h_masks = None
    mon = FluxMonitorTM(grid, [0.6, 0.8, 1.0], x_normal=True, idx=38, lo=7, hi=NY - 7)
    st = init_tm_state(grid, n_poles=2 if case == "mag" else 0, n_mpoles=2)
    return g, m, src, mon, st


@pytest.mark.parametrize("case", ["mag", "magh"])
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
@pytest.mark.parametrize("chunk", [0, 70])
def test_magnetic_poles_match_reference(case, tag, dtype, chunk):
    from meep_torch_amd.core.run import run_tm
    g, m, src, mon, st = _setup(case, tag, dtype)
    print(f"[mag_tm] {case} {tag} chunk={chunk}", flush=True)
    st, (acc,) = run_tm(st, m, N, source=src, monitors=[mon], checkpoint_every=chunk)
    for f in NAMES + (("P_cur", "P_prev") if case == "mag" else ()):
        got, ref = getattr(st, f).cpu(), torch.from_numpy(g[f"{case}_{f}"])
        assert torch.equal(got, ref), (case, f, float((got - ref).abs().max()))
    ra = torch.from_numpy(g[f"{case}_acc"])
    tol = 1e-4 if dtype == torch.float32 else 1e-12
    assert float((acc.cpu() - ra).norm() / ra.norm()) <= tol
    assert float(st.My_cur.abs().max()) > 0


@pytest.mark.parametrize("case", ["mag", "magh"])
def test_magnetic_pole_gradients_match_oracle_autograd(case):
    """The TM adjoint transposes the magnetic poles' B / M recurrence
    (H' = B' - sum M', M' = (mc1 M + mc2 M_prev) + mc3 H, B' = B +- dt curl):
    the eps gradient of a flux objective equals CPU autograd through the
    oracle's steps (FP64), with electric poles and chi3 in the same cells
    (mag) and under a TFSF Hy current (magh); dropping the M recurrence's
    H term changes the gradient by far more than the bar.  A batch runs the
    variant through the same path (its gradient equals run_tm's)."""
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.run import run_tm
    from oracle import yee2d
    g, m, src, mon, st = _setup(case, "f64", torch.float64)
    n = 160
    eps = m.eps_inf.clone().requires_grad_(True)
    _, (acc,) = run_tm(st, m._replace(eps_inf=eps), n, source=src, monitors=[mon], checkpoint_every=60)
    w = torch.linspace(0.5, 1.5, acc.numel(), dtype=acc.dtype, device=acc.device).reshape(acc.shape)
    (ge,) = torch.autograd.grad((acc * w).sum(), [eps])
    t = lambda k: torch.from_numpy(np.ascontiguousarray(g[k]))
    torch.set_num_threads(8)

    def oracle_grad(mpoles):
        e = t("eps").clone().requires_grad_(True)
        kw = dict(m_poles=mpoles, m_pole_masks=[t("mm1"), t("mm2")])
        if case == "mag":
            kw.update(poles=POLES2, pole_masks=[t("m1"), t("m2")], chi3=t("chi3"))
        c = yee2d.tm_coeffs(NX, NY, 0.05, e, 6, dtype=torch.float64, **kw)
        ost = yee2d.tm_state(NX, NY, torch.float64, n_poles=2 if case == "mag" else 0, n_mpoles=2)
        mons = [{"type": "flux", "kind": "tm_x", "idx": 38, "lo": 7, "hi": NY - 7, "freqs": [0.6, 0.8, 1.0]}]
        hk = {} if case == "mag" else dict(hamps=g["magh_h_amp"][:n], hmasks={"hy_mask": t("magh_hy_mask")})
        _, (oacc,) = yee2d.run2d("tm", ost, c, n, amps=g["mag_amp"][:n], mask=t("mag_mask"), monitors=mons, **hk)
        (gr,) = torch.autograd.grad((oacc * w.cpu().reshape(oacc.shape)).sum(), e)
        return gr

    g_ref = oracle_grad(MPOLES2)
    err = float((ge.cpu() - g_ref).norm() / g_ref.norm())
    g_off = oracle_grad([(fn, gam, 0.0) for fn, gam, _ in MPOLES2])   # no H drive of the M stacks
    off = float((g_off - g_ref).norm() / g_ref.norm())
    print(f"\n[magnetic pole grad {case}] rel-L2 {err:.2e} (without the M drive {off:.2e})")
    assert err <= 1e-9 and off > 1e-4
    bt = BatchedTM([(m, src, [mon])], n, checkpoint_every=60)
    bt.forward([m.eps_inf])
    (gb,) = bt.gradient([[w]])
    assert torch.equal(gb, ge)
