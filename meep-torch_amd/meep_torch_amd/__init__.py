"""meep_torch_amd: MI355X-native FDTD time-stepping engine behind the
meep_torch API (Grid / build_*_coeffs / run_tm / run_te / run_3d / monitors /
sources / Simulation).  Module layout mirrors meep_torch so callers switch by
import name; the stepping runs in hand-written HIP kernels (libmeepfdtd.so)."""
__version__ = "0.1.0"
