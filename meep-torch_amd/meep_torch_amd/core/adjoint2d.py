"""Hand-written time-reversed adjoint for run_tm (placeholder until the
adjoint kernels land)."""


def run_with_adjoint(mode, state, coeffs, n_steps, source, monitors, checkpoint_every):
    raise NotImplementedError("gradient through run_tm/run_te: adjoint kernels not built yet")
