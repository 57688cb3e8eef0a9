"""mphip: MI355X (gfx950) backend for MagnetoPhotonic.jl's coupled 3-D time step.

Host-side mirror of the reference's FDTDState / step! API over the C-ABI library
libmphip.so (hand-written HIP kernels).  Grid, CPML, pole, parameter and geometry
preparation are host numpy (one-time setup); every per-step operation runs on the GPU.
"""
from .grid import (C0, EPS0, MU0, Axis1D, FDTDParams, Grid3D, cfl_dt, cfl_dt_cells, cfl_dt_spacing,
                   graded_axis, not_gate_reference_grid, uniform_axis, uniform_axis_n)
from .models import FerrimagnetParameters, ferrimagnet_parameters
from .sources import (CarrierGaussianPulse, ContinuousSource, GaussianPulse, ModeSource, PlaneSource,
                      SoftSource)
from .geometry import Box, Material, Scene, Sphere, rasterize
from .spec import StateSpec, build_state_spec

__all__ = [
    "C0", "EPS0", "MU0", "Axis1D", "FDTDParams", "Grid3D", "cfl_dt", "cfl_dt_cells", "cfl_dt_spacing",
    "graded_axis", "not_gate_reference_grid", "uniform_axis", "uniform_axis_n",
    "FerrimagnetParameters", "ferrimagnet_parameters", "CarrierGaussianPulse", "ContinuousSource",
    "GaussianPulse", "ModeSource", "PlaneSource", "SoftSource", "Box", "Material", "Scene", "Sphere",
    "rasterize", "StateSpec", "build_state_spec", "FDTDState",
]


def __getattr__(name):
    # The device state loads libmphip lazily so host-only setup imports without a GPU.
    if name == "FDTDState":
        from .state import FDTDState
        return FDTDState
    raise AttributeError(name)
