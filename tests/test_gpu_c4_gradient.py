"""C4 at full size in FP64 (24-cell CPML, 4000 steps, checkpoints every 200):
density gradients -- batched forward + hand-written batched adjoint, then
autograd through the tanh projection and conic filter of each 128 x 128
density -- against central finite differences (h = 1e-4, as the reference's
own FD checks).

* the first two seeded variants (954^2 and 1426^2) in one batch: 8 seeded
  pixels of variant 1's 20-frequency output flux; variant 0's density must
  get an exactly zero gradient;
* the first 32 seeded variants in one batch, FOM = the sum of their fluxes
  (C4's FOM over those variants): one seeded pixel per variant, all 32
  perturbed in the same batched forward (each variant's flux depends on its
  own density only, so one +h and one -h forward give 32 independent
  central differences)."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def test_c4_density_gradient_central_fd():
    from meep_torch_amd import synth
    from meep_torch_amd.adjoint.filters import conic_filter, tanh_projection
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    dt = torch.float64
    u = synth.splitmix_uniform(64, 4)
    sizes = [512 + int(x * 1025) for x in u][:2]
    assert sizes == [954, 1426]
    npml, fr = 24, [0.4 + 0.2 * i / 19 for i in range(20)]
    probs, bases, mons = [], [], []
    for n in sizes:
        g = Grid(nx=n, ny=n, dx=0.05, dtype=dt, device="cuda")
        base = torch.full((n, n), 2.1, dtype=dt, device="cuda")
        m = build_tm_coeffs(g, eps_inf=base, npml=npml)
        src = LineSource(g, ix=npml + 16, j0=npml, j1=n - npml, profile=GaussianPulse(0.5, 0.2))
        mon = FluxMonitorTM(g, fr, x_normal=True, idx=n - npml - 16, lo=npml, hi=n - npml)
        probs.append((m, src, [mon]))
        bases.append(base)
        mons.append(mon)
    rho = [torch.tensor(synth.splitmix_uniform(128 * 128, 100 + v).reshape(128, 128), dtype=dt, device="cuda")
           for v in range(2)]
    bt = BatchedTM(probs, 4000, checkpoint_every=200)

    def eps_of(v, r):
        n = sizes[v]
        e = bases[v].clone()
        o = (n - 128) // 2
        e[o:o + 128, o:o + 128] = 2.1 + 10.0 * tanh_projection(conic_filter(r, 0.2, 0.05), 8.0, 0.5)
        return e

    def fom(out):
        return mons[1].flux(out[1][0], 0.05).sum()

    rs = [r.clone().requires_grad_(True) for r in rho]
    J = fom(bt.apply([eps_of(v, r) for v, r in enumerate(rs)]))
    J.backward()
    grad = rs[1].grad.detach()
    assert float(rs[0].grad.abs().max()) == 0.0
    gmax = float(grad.abs().max())
    assert gmax > 0
    pix = [(int(a * 128), int(b * 128)) for a, b in zip(synth.splitmix_uniform(8, 7), synth.splitmix_uniform(8, 8))]
    h = 1e-4
    worst = 0.0
    with torch.no_grad():
        for (i, j) in pix:
            rp, rm = rho[1].clone(), rho[1].clone()
            rp[i, j] += h
            rm[i, j] -= h
            fp = float(fom(bt.forward([eps_of(0, rho[0]), eps_of(1, rp)])))
            fm = float(fom(bt.forward([eps_of(0, rho[0]), eps_of(1, rm)])))
            fd = (fp - fm) / (2 * h)
            g = float(grad[i, j])
            err = abs(g - fd) / max(abs(fd), 1e-2 * gmax)
            worst = max(worst, err)
            assert err <= 1e-4, ((i, j), g, fd, err)
    print(f"\n[c4 fd] J = {float(J.detach()):.6e}, max |dJ/drho| = {gmax:.3e}, worst relative FD error {worst:.2e}")


def _c4_problems(n_var, dt):
    from meep_torch_amd import synth
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    u = synth.splitmix_uniform(64, 4)
    sizes = [512 + int(x * 1025) for x in u][:n_var]
    npml, fr = 24, [0.4 + 0.2 * i / 19 for i in range(20)]
    probs, bases, mons = [], [], []
    for n in sizes:
        g = Grid(nx=n, ny=n, dx=0.05, dtype=dt, device="cuda")
        base = torch.full((n, n), 2.1, dtype=dt, device="cuda")
        m = build_tm_coeffs(g, eps_inf=base, npml=npml)
        src = LineSource(g, ix=npml + 16, j0=npml, j1=n - npml, profile=GaussianPulse(0.5, 0.2))
        mon = FluxMonitorTM(g, fr, x_normal=True, idx=n - npml - 16, lo=npml, hi=n - npml)
        probs.append((m, src, [mon]))
        bases.append(base)
        mons.append(mon)
    return sizes, probs, bases, mons


def test_c4_32_variants_one_pixel_each_central_fd():
    from meep_torch_amd import synth
    from meep_torch_amd.adjoint.filters import conic_filter, tanh_projection
    from meep_torch_amd.batch import BatchedTM
    dt, nv = torch.float64, 32
    sizes, probs, bases, mons = _c4_problems(nv, dt)
    rho = [torch.tensor(synth.splitmix_uniform(128 * 128, 100 + v).reshape(128, 128), dtype=dt, device="cuda")
           for v in range(nv)]
    bt = BatchedTM(probs, 4000, checkpoint_every=200)

    def eps_of(v, r):
        n = sizes[v]
        e = bases[v].clone()
        o = (n - 128) // 2
        e[o:o + 128, o:o + 128] = 2.1 + 10.0 * tanh_projection(conic_filter(r, 0.2, 0.05), 8.0, 0.5)
        return e

    def fluxes(out):
        return [mons[v].flux(o[0], 0.05).sum() for v, o in enumerate(out)]

    rs = [r.clone().requires_grad_(True) for r in rho]
    J = sum(fluxes(bt.apply([eps_of(v, r) for v, r in enumerate(rs)])))
    J.backward()
    pi = synth.splitmix_uniform(nv, 9)
    pj = synth.splitmix_uniform(nv, 10)
    pix = [(int(a * 128), int(b * 128)) for a, b in zip(pi, pj)]
    h = 1e-4
    with torch.no_grad():
        rp, rm = [r.clone() for r in rho], [r.clone() for r in rho]
        for v, (i, j) in enumerate(pix):
            rp[v][i, j] += h
            rm[v][i, j] -= h
        fp = [float(x) for x in fluxes(bt.forward([eps_of(v, r) for v, r in enumerate(rp)]))]
        fm = [float(x) for x in fluxes(bt.forward([eps_of(v, r) for v, r in enumerate(rm)]))]
    worst = 0.0
    for v, (i, j) in enumerate(pix):
        gv = rs[v].grad.detach()
        gmax = float(gv.abs().max())
        assert gmax > 0, v
        fd = (fp[v] - fm[v]) / (2 * h)
        g = float(gv[i, j])
        err = abs(g - fd) / max(abs(fd), 1e-2 * gmax)
        worst = max(worst, err)
        assert err <= 1e-4, (v, sizes[v], (i, j), g, fd, err)
    print(f"\n[c4 fd, {nv} variants N {min(sizes)}..{max(sizes)}] J = {float(J.detach()):.6e}, "
          f"worst relative FD error over {nv} pixels {worst:.2e}")
