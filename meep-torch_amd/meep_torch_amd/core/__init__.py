"""Engine core: the meep_torch.core API surface on the HIP stepper."""
