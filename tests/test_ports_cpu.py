"""Host side of the eigenmode ports (meep_torch/ports/, modes/fd1d.py):
the 1D mode solver against closed forms and the analytic slab dispersion,
the modes' bi-orthogonality, port validation and the placement of the TFSF
source sheets."""
import math

import numpy as np
import pytest
import torch


def _slab(n_core, n_clad, d, dx, n):
    y = (np.arange(n) + 0.5) * dx
    return torch.tensor(np.where(np.abs(y - n * dx / 2) < d / 2, n_core ** 2, n_clad ** 2), dtype=torch.float64)


def _bisect(g, lo, hi):
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(lo) * g(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def _analytic_even(n1, n2, d, lam, te):
    """Fundamental even mode of a symmetric slab: kappa tan(kappa d/2) = r gamma
    with r = 1 for E along the interfaces (the 'tm' / Ez operator) and
    r = n1^2 / n2^2 for H along them (the 'te' / Hz operator); the root lies
    on the first branch, kappa d/2 in (0, pi/2)."""
    k0 = 2 * math.pi / lam
    r = (n1 / n2) ** 2 if te else 1.0

    def g(neff):
        kap = k0 * math.sqrt(n1 ** 2 - neff ** 2)
        gam = k0 * math.sqrt(neff ** 2 - n2 ** 2)
        return kap * math.tan(kap * d / 2) - r * gam
    pole = n1 ** 2 - (math.pi / (k0 * d)) ** 2
    lo = math.sqrt(pole) + 1e-9 if pole > n2 ** 2 else n2 + 1e-9
    return _bisect(g, lo, n1 - 1e-9)


@pytest.mark.parametrize("pol", ["tm", "te"])
def test_uniform_line_closed_form(pol):
    """Uniform eps between PEC walls: A = D2 / k0^2 + eps I for both
    operators, eigenvalues eps - 4 s sin^2(pi m / (2 (n + 1))), eigenvectors
    sin(pi m (i + 1) / (n + 1))."""
    from meep_torch_amd.ports import solve_1d
    n, dx, f, eps = 40, 0.1, 0.3, 2.25
    m = solve_1d(torch.full((n,), eps, dtype=torch.float64), dx, f, pol, k=3)
    s = 1.0 / (2 * math.pi * f * dx) ** 2
    i = np.arange(n)
    for q in range(3):
        lam = eps - 4 * s * math.sin(math.pi * (q + 1) / (2 * (n + 1))) ** 2
        assert float(m.neff[q]) == pytest.approx(math.sqrt(lam), rel=1e-12)
        v = np.sin(math.pi * (q + 1) * (i + 1) / (n + 1))
        v /= np.linalg.norm(v)
        v *= np.sign(v[np.argmax(np.abs(v))])
        assert np.allclose(m.profiles[q].numpy(), v, atol=1e-10)


def test_pmc_low_end_closed_form():
    """Neumann ghost at the low end, Dirichlet at the high end: eigenvectors
    cos(pi (m - 1/2) (i + 1/2) / (n + 1/2)), eigenvalues
    eps - 4 s sin^2(pi (m - 1/2) / (2 n + 1))."""
    from meep_torch_amd.ports import solve_1d
    n, dx, f, eps = 30, 0.1, 0.3, 4.0
    m = solve_1d(torch.full((n,), eps, dtype=torch.float64), dx, f, "tm", k=2, bc="pmc_lo")
    s = 1.0 / (2 * math.pi * f * dx) ** 2
    i = np.arange(n)
    for q in range(2):
        lam = eps - 4 * s * math.sin(math.pi * (q + 0.5) / (2 * n + 1)) ** 2
        assert float(m.neff[q]) == pytest.approx(math.sqrt(lam), rel=1e-12)
        v = np.cos(math.pi * (q + 0.5) * (i + 0.5) / (n + 0.5))
        v /= np.linalg.norm(v)
        v *= np.sign(v[np.argmax(np.abs(v))])
        assert np.allclose(m.profiles[q].numpy(), v, atol=1e-10)


@pytest.mark.parametrize("pol,tol", [("tm", 2e-4), ("te", 2e-3)])
def test_slab_matches_analytic_dispersion(pol, tol):
    """Si / SiO2 slab, 0.5 um, 1.55 um wavelength, dx = 5 nm."""
    from meep_torch_amd.ports import solve_1d
    eps = _slab(3.4777, 1.444, 0.5, 0.005, 1200)
    m = solve_1d(eps, 0.005, 1 / 1.55, pol, k=1)
    ref = _analytic_even(3.4777, 1.444, 0.5, 1.55, te=(pol == "te"))
    assert abs(float(m.neff[0]) - ref) < tol, (float(m.neff[0]), ref)
    assert 500 < int(torch.argmax(m.profiles[0].abs())) < 700


def test_te_pencil_equals_the_nonsymmetric_operator():
    """The TE solve's symmetric pencil has the eigenpairs of
    A = diag(eps) (d/dy (1/eps_edge) d/dy / k0^2 + I), built densely here
    and solved with a general eigensolver."""
    from meep_torch_amd.ports import solve_1d
    eps = _slab(3.0, 1.2, 1.0, 0.05, 80).numpy()
    n, dx, f = eps.size, 0.05, 0.4
    k0 = 2 * math.pi * f
    eh = np.concatenate([[eps[0]], 0.5 * (eps[:-1] + eps[1:]), [eps[-1]]])
    Df = (np.eye(n + 1, n) - np.eye(n + 1, n, -1)) / dx
    Db = (np.eye(n, n + 1, 1) - np.eye(n, n + 1)) / dx
    A = np.diag(eps) @ (Db @ np.diag(1 / eh) @ Df / k0 ** 2 + np.eye(n))
    w, V = np.linalg.eig(A)
    top = np.argsort(w.real)[::-1][:3]
    m = solve_1d(torch.from_numpy(eps), dx, f, "te", k=3)
    for q, c in enumerate(top):
        assert float(m.neff[q]) == pytest.approx(math.sqrt(w[c].real), rel=1e-10)
        v = V[:, c].real / np.linalg.norm(V[:, c].real)
        assert abs(abs(float(np.dot(v, m.profiles[q].numpy()))) - 1) < 1e-9


def test_modes_are_biorthonormal_under_the_decomposition_norm():
    """For TM, N_mn = sum(E_m neff_n E_n) dx is diagonal (a 3 um eps-12 core
    guides three modes)."""
    from meep_torch_amd.ports import solve_1d
    eps = torch.ones(60, dtype=torch.float64)
    eps[15:45] = 12.0
    m = solve_1d(eps, 0.1, 0.15, "tm", k=3)
    for a in range(3):
        na = float(m.neff[a]) * float((m.profiles[a] ** 2).sum()) * 0.1
        for b in range(3):
            o = float((m.profiles[a] * float(m.neff[b]) * m.profiles[b]).sum()) * 0.1
            assert abs(o / na - (1.0 if a == b else 0.0)) < 1e-10


def test_validate_ports_and_source_placement():
    from meep_torch_amd.core.coeffs import build_te_coeffs, build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.ports import ModeSourceTE, ModeSourceTM, Port, validate_ports
    g = Grid(nx=160, ny=80, dx=0.1, dtype=torch.float64, device="cpu")
    for bad, msg in ((Port("z", 80, 20, 60, 1), "axis"), (Port("x", 80, 20, 60, 0), "direction"),
                     (Port("x", 11, 20, 60, 1), "PML"), (Port("x", 80, 5, 60, 1), "PML"),
                     (Port("x", 80, 40, 40, 1), "empty")):
        with pytest.raises(ValueError, match=msg):
            validate_ports([bad], g, 10)
    eps = torch.ones(160, 80, dtype=torch.float64)
    eps[:, 35:45] = 12.0
    c = build_tm_coeffs(g, eps_inf=eps, npml=10)
    for d in (1, -1):
        s = ModeSourceTM(g, c, Port("x", 80, 20, 60, d), GaussianPulse(0.15, 0.1), 0.15)
        j = 80 - 3 * d
        assert torch.nonzero(s.mask.abs().sum(1)).flatten().tolist() == [j]
        hy = s.h_masks["src_hy_mask"]
        assert torch.nonzero(hy.abs().sum(1)).flatten().tolist() == [j - 1 if d > 0 else j]
        assert 1.0 < s.neff < math.sqrt(12.0) and s.delay == pytest.approx(s.neff * 0.05)
        # the magnetic sheet fires neff dx / 2 later, with the TM sign
        assert s.h_amplitude(40) == pytest.approx(-GaussianPulse(0.15, 0.1).amplitude(40 * g.dt - s.delay))
    with pytest.raises(ValueError, match="guided"):
        ModeSourceTM(g, build_tm_coeffs(g, eps_inf=1.0, npml=10), Port("x", 80, 20, 60, 1),
                     GaussianPulse(0.15, 0.1), 0.15)
    ct = build_te_coeffs(g, eps_inf=eps, npml=10)
    s = ModeSourceTE(g, ct, Port("x", 80, 20, 60, 1), GaussianPulse(0.15, 0.1), 0.15)
    assert torch.nonzero(s.mask.abs().sum(1)).flatten().tolist() == [68]
    assert torch.nonzero(s.h_masks["src_hz_mask"].abs().sum(1)).flatten().tolist() == [67]
    with pytest.raises(NotImplementedError):
        ModeSourceTE(g, ct, Port("y", 80, 20, 60, 1), GaussianPulse(0.15, 0.1), 0.15)
