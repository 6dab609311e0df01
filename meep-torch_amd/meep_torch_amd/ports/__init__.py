"""Eigenmode ports over the HIP engine (meep_torch/ports/): 1D port-line mode
solves, one-way TFSF mode sources, modal decomposition of flux lines and the
per-port S-matrix driver (2D)."""
from .mode_monitor import decompose, dft_fields
from .mode_source import ModeSourceTE, ModeSourceTM
from .modes1d import Modes1D, solve_1d
from .port import SRC_OFFSET, TE_SRC_OFFSET, Port, validate_ports
from .smatrix import SMatrix, compute_smatrix

__all__ = ["Port", "validate_ports", "SRC_OFFSET", "TE_SRC_OFFSET", "Modes1D", "solve_1d", "ModeSourceTM",
           "ModeSourceTE", "decompose", "dft_fields", "SMatrix", "compute_smatrix"]
