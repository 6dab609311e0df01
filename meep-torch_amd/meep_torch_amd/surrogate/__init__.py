"""Learned surrogate for inverse design: a spectral-convolution neural
operator (fno.SpectralOperator) trained on transmission spectra computed by
the HIP FDTD engine (problem.TransmissionProblem, train.build_surrogate),
used inside a design loop that verifies it against FDTD and falls back to
FDTD gradients when it disagrees (design.surrogate_loop)."""
