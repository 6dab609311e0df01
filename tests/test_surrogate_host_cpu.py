"""Host-side pieces of the surrogate workflow that need no GPU: the
parameter layout (4,225,970 parameters at the default size), seeded design
sampling (deterministic, in [0, 1], all three families present), the
spectral layers' mode set, and that the GPU entry points refuse CPU
tensors instead of falling back."""
import numpy as np
import pytest
import torch


def test_default_parameter_count_and_modes():
    from meep_torch_amd.surrogate.fno import mode_list
    C, L, Mo, H, n_out, in_ch = 64, 4, 16 * 8, 128, 50, 3
    total = C * in_ch + C + L * (2 * C * C * Mo + C * C + C) + H * C + H + n_out * H + n_out
    assert total == 4225970
    m = mode_list(128, 16, 8)
    assert len(m) == 128 and len(set(m)) == 128
    assert {kx for kx, _ in m} == set(range(8)) | set(range(120, 128))
    with pytest.raises(ValueError):
        mode_list(16, 4, 9)      # Nyquist column excluded


def test_sample_designs_seeded_and_bounded():
    from meep_torch_amd.surrogate.problem import sample_designs
    a = sample_designs(300, 5, device="cpu")
    b = sample_designs(300, 5, device="cpu")
    c = sample_designs(300, 6, device="cpu")
    assert torch.equal(a, b) and not torch.equal(a, c)
    assert float(a.min()) >= 0.0 and float(a.max()) <= 1.0
    # channel-like samples: strong contrast between a horizontal band and the rest
    band = (a[:, :, 54:74].mean((1, 2)) - a.mean((1, 2))).numpy()
    assert (band > 0.15).sum() > 30 and (np.abs(band) < 0.05).sum() > 30


def test_bases_match_oracle_definition():
    from meep_torch_amd.surrogate.fno import dft_bases
    from oracle import fno as O
    F, G = dft_bases(24, 6, 4)
    Fo, Go = O.bases(24, 6, 4)
    assert np.abs(F - Fo).max() < 1e-14 and np.abs(G - Go).max() < 1e-14


def test_gpu_entry_points_refuse_cpu_tensors():
    from meep_torch_amd.surrogate import _nn
    a = torch.zeros(4, 4, dtype=torch.bfloat16)
    c = torch.zeros(4, 4)
    with pytest.raises(RuntimeError):
        _nn.gemm(a, (4, 1, 0, 0), a, (4, 1, 0, 0), c, (4, 1, 0), 4, 4, 4)
