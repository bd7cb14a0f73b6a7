"""MI355X-native M1 proton energy march with the reference solver API."""

from .beam import Beam
from .config import DeviceOptions, Numerics, OutputConfig, Physics, References, SolverConfig
from .enums import (
    Backend,
    CoefficientInterpolation,
    EnergyScheme,
    Material,
    Precision,
    Reference,
    Scattering,
    SpatialScheme,
    WaveSpeeds,
)
from .geometry import Box, CTScan, Geometry, HUPartition

__all__ = [
    "Solver", "SolverConfig",
    "Geometry", "Box", "CTScan", "HUPartition", "Beam",
    "Physics", "Numerics", "References", "OutputConfig", "DeviceOptions",
    "Backend", "Material", "SpatialScheme", "EnergyScheme", "WaveSpeeds",
    "Reference", "Scattering", "CoefficientInterpolation", "Precision",
]


def __getattr__(name):
    # The solver pulls in the native march library; load it on first use.
    if name == "Solver":
        from .solver import Solver
        return Solver
    raise AttributeError(name)
