"""Design chain (meep_torch/adjoint): filters, projection, MaterialGrid,
DesignRegion, OptimizationProblem."""
from .filters import conic_filter, conic_filter3d, gaussian_filter, tanh_projection  # noqa: F401
