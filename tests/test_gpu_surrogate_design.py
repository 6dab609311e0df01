"""What the trained spectral operator gains over the method without it, on
the transmission design problem (reduced dataset so the test runs in about
a minute): against an untrained network of the same architecture, the
trained one (a) predicts held-out FDTD transmission spectra with lower
error and (b) gives density gradients of the mean transmission that point
along the FDTD adjoint gradient (cosine similarity) -- the quantity the
design loop consumes.  Also checks that the surrogate-assisted loop runs
its FDTD verification cadence and leaves a design whose FDTD-measured FOM
improved on the start."""
import pytest
import torch

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


@pytest.fixture(scope="module")
def trained():
    from meep_torch_amd.surrogate.fno import SpectralOperator
    from meep_torch_amd.surrogate.problem import TransmissionProblem, sample_designs
    from meep_torch_amd.surrogate.train import generate_dataset, train
    pb = TransmissionProblem()
    rho, T = generate_dataset(pb, 1536, seed=7)
    theta_v = sample_designs(96, 99, pb.d, pb.device)
    with torch.no_grad():
        rho_v = pb.project(theta_v)
    T_v = pb.transmission(rho_v)
    net = SpectralOperator(n=pb.d, n_out=len(pb.freqs), seed=1)
    fresh = SpectralOperator(n=pb.d, n_out=len(pb.freqs), seed=1)
    train(net, rho, T, epochs=10, seed=3)
    return pb, net, fresh, (rho, T), (theta_v, rho_v, T_v)


def _err(net, rho, T):
    from meep_torch_amd.surrogate.train import spectrum_error
    with torch.no_grad():
        return float(spectrum_error(net.forward(rho), T).median())


def _grad(pb, net, theta):
    from meep_torch_amd.surrogate.fno import predict
    th = theta.detach().clone().requires_grad_(True)
    predict(net, pb.project(th)[None])[0].mean().backward()
    return th.grad.reshape(-1)


def test_trained_network_beats_untrained(trained):
    pb, net, fresh, _, (theta_v, rho_v, T_v) = trained
    e_tr, e_un = _err(net, rho_v, T_v), _err(fresh, rho_v, T_v)
    cos_tr, cos_un = [], []
    for t in range(4):
        _, _, g_fdtd = pb.solve_with_gradient(theta_v[t])
        g = g_fdtd.reshape(-1)
        cos_tr.append(float(torch.nn.functional.cosine_similarity(_grad(pb, net, theta_v[t]), g, dim=0)))
        cos_un.append(float(torch.nn.functional.cosine_similarity(_grad(pb, fresh, theta_v[t]), g, dim=0)))
    print(f"\n[surrogate] held-out median spectrum error: trained {e_tr:.3f}, untrained {e_un:.3f}; "
          f"gradient cosine vs FDTD adjoint: trained {[round(c, 3) for c in cos_tr]}, "
          f"untrained {[round(c, 3) for c in cos_un]}")
    assert e_tr < 0.8 * e_un
    assert sum(cos_tr) / 4 > sum(cos_un) / 4


def test_surrogate_loop_verifies_and_improves(trained):
    from meep_torch_amd.surrogate import design as D
    pb, net, _, data, _ = trained
    theta0 = D.start_design(pb, seed=5)
    _, fom0, _ = pb.solve_with_gradient(theta0)
    res = D.surrogate_loop(pb, net, iters=30, theta0=theta0, data=data)
    print(f"\n[surrogate loop] start FOM {fom0:.4f} -> {res['final_fom']:.4f}; "
          f"{res['verifications']} verifications ({res['failed']} failed), {res['fdtd_iters']} FDTD / "
          f"{res['surrogate_iters']} surrogate iterations; errors {[round(e, 3) for e in res['errors']]}")
    assert res["verifications"] == 3
    assert res["fdtd_iters"] + res["surrogate_iters"] == 30
    assert res["final_fom"] > fom0
