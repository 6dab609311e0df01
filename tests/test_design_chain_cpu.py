"""Design chain host pieces against the reference's golden vectors
(tests/golden/filters.npz, written from meep_torch/adjoint/filters.py and
problem.py): conic and Gaussian filters, tanh projection and DesignRegion
painting (two upsampling ratios) bit-exact in float64; plus the
OptimizationProblem plumbing with a CPU objective (gradient by central FD)."""
import numpy as np
import torch

from conftest import golden


def test_filters_and_paint_bitexact_vs_golden():
    from meep_torch_amd.adjoint.filters import conic_filter, gaussian_filter, tanh_projection
    from meep_torch_amd.adjoint.problem import DesignRegion, MaterialGrid
    g = golden("filters")
    rho = torch.tensor(g["rho"])
    cf = conic_filter(rho, 0.2, 0.05)
    assert np.array_equal(cf.numpy(), g["conic"])
    assert np.array_equal(gaussian_filter(rho, 0.1, 0.05).numpy(), g["gauss"])
    tp = tanh_projection(cf, 8.0, 0.5)
    assert np.array_equal(tp.numpy(), g["tanh"])
    base = torch.ones(60, 40, dtype=torch.float64)
    dr = DesignRegion(MaterialGrid((32, 24), 2.1, 12.1), ix0=10, ix1=42, iy0=5, iy1=29)
    assert np.array_equal(dr.paint(base, tp).numpy(), g["painted"])
    dr2 = DesignRegion(MaterialGrid((32, 24), 2.1, 12.1), ix0=4, ix1=52, iy0=3, iy1=39)
    assert np.array_equal(dr2.paint(base, tp).numpy(), g["painted2"])


def test_optimization_problem_gradient_fd():
    from meep_torch_amd.adjoint.problem import DesignRegion, MaterialGrid, OptimizationProblem
    torch.manual_seed(0)
    w = torch.randn(20, 16, dtype=torch.float64)
    base = torch.ones(20, 16, dtype=torch.float64)
    dr = DesignRegion(MaterialGrid((8, 6), 2.0, 10.0), ix0=4, ix1=16, iy0=2, iy1=14)
    opt = OptimizationProblem(lambda eps: (w * eps).sum() + (eps ** 2).mean(), base, dr,
                              filter_radius=0.15, dx=0.05, beta=4.0)
    opt.rho = torch.rand(8, 6, dtype=torch.float64)
    J, grad = opt()
    for (i, j) in [(0, 0), (3, 2), (7, 5)]:
        fd = (opt.eval_at_delta(i, j, 1e-6) - opt.eval_at_delta(i, j, -1e-6)) / 2e-6
        assert abs(grad[i, j] - fd) < 1e-6 * max(1.0, abs(fd))
