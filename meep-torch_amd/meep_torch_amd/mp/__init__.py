"""meep-compatible facade (subset) over the HIP FDTD engine: Simulation,
Block/Ellipsoid/Cylinder/Sphere/Prism rasterization, media, PML, Gaussian/continuous line sources, flux and
DFT-field monitors, step functions and stop conditions (meep_torch/mp)."""
from .geometry import (PML, Absorber, Block, Cylinder, DrudeSusceptibility, Ellipsoid, GeometricObject,
                       LorentzianSusceptibility, Medium, Prism, Sphere)
from .run_control import (FluxData, after_sources, after_time, at_beginning, at_end, at_every, at_time,
                          before_time, during_sources, stop_after_walltime, stop_when_fields_decayed)
from .simulation import (DftFields, DftFlux, EigCoeffsResult, FluxRegion, Simulation, Volume, get_flux_freqs,
                         get_fluxes)
from .sources import AUTOMATIC, ContinuousSource, EigenModeSource, GaussianSource, Source
from .symmetry import Identity, Mirror, Rotate2, Rotate4, Symmetry
from .vec import (ALL, Dielectric, Dz, Ex, Ey, Ez, High, Hx, Hy, Hz, Low, Vector3, X, Y, Z, inf)

__all__ = ["PML", "Absorber", "Block", "Cylinder", "Ellipsoid", "Prism", "Sphere", "DrudeSusceptibility", "GeometricObject", "LorentzianSusceptibility", "Medium", "DftFields", "DftFlux", "FluxRegion",
           "Simulation", "Volume", "get_flux_freqs", "get_fluxes", "ContinuousSource", "GaussianSource",
           "Source", "ALL", "Dielectric", "Dz", "Ex", "Ey", "Ez", "High", "Hx", "Hy", "Hz", "Low", "Vector3",
           "X", "Y", "Z", "inf", "FluxData", "after_sources", "after_time", "at_beginning", "at_end",
           "at_every", "at_time", "before_time", "during_sources", "stop_after_walltime",
           "stop_when_fields_decayed", "AUTOMATIC", "EigenModeSource", "EigCoeffsResult", "Identity", "Mirror",
           "Rotate2", "Rotate4", "Symmetry"]
