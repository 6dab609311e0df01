"""Grid (meep_torch/core/grid.py:33-59): nx, ny, dx, Courant, dtype, device.

Same validation and dt = courant * dx.  The default device is read from
MEEP_TORCH_DEVICE as in the reference (core/grid.py:10-30), but defaults to
the ROCm device "cuda": the engine has no CPU stepper."""
import math
import os
from dataclasses import dataclass, field

import torch


def default_device() -> str:
    return os.environ.get("MEEP_TORCH_DEVICE", "cuda")


@dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    dx: float
    courant: float = 0.5
    dtype: torch.dtype = torch.float64
    device: str = field(default_factory=default_device)

    def __post_init__(self):
        if self.nx < 3:
            raise ValueError(f"nx must be >= 3, got {self.nx}")
        if self.ny < 3:
            raise ValueError(f"ny must be >= 3, got {self.ny}")
        if self.dx <= 0:
            raise ValueError(f"dx must be > 0, got {self.dx}")
        if not 0 < self.courant <= 1 / math.sqrt(2):
            raise ValueError(
                f"courant must be in (0, 1/sqrt(2)] for 2D stability, got {self.courant}")

    @property
    def dt(self) -> float:
        return self.courant * self.dx

    def zeros(self, *shape) -> torch.Tensor:
        return torch.zeros(*shape, dtype=self.dtype, device=self.device)
