"""Enumerations of the solver API (mirrors ``proton_solver/enums.py:4-51``).

``Material`` carries two members the reference enum does not list, AIR and
MUSCLE_SKELETAL: the reference ships their tables (``data/materials.csv:2,6``)
but no enum entry.  All reference members keep their names and values.
"""

from enum import Enum, auto


class Backend(Enum):
    AUTO = auto()
    CPU = auto()
    GPU = auto()


class Material(Enum):
    WATER = "G4_WATER"
    BONE_COMPACT = "G4_BONE_COMPACT_ICRU"
    LUNG = "G4_LUNG_ICRP"
    AIR = "G4_AIR"
    MUSCLE_SKELETAL = "G4_MUSCLE_SKELETAL_ICRP"


class SpatialScheme(Enum):
    FIRST_ORDER = auto()
    MUSCL = auto()
    WENO3 = auto()
    WENO_Z = auto()


class EnergyScheme(Enum):
    FORWARD_EULER = auto()
    BACKWARD_EULER = auto()
    TRAPEZOIDAL = auto()
    SSPRK2 = auto()
    SSPRK3 = auto()
    IF_SSPRK3 = auto()


class WaveSpeeds(Enum):
    EXACT = auto()
    UNIT = auto()


class Reference(Enum):
    NONE = auto()
    ANALYTICAL = auto()
    TOPAS = auto()
    BFP_MC = auto()
    SELF = auto()


class CoefficientInterpolation(Enum):
    """Interpolation of the tabulated scattering and removal coefficients."""

    LINEAR = auto()
    PCHIP = auto()


class Scattering(Enum):
    OFF = auto()
    ROSSI = auto()
    WENTZEL = auto()


class Precision(Enum):
    """Arithmetic of the device march (no reference counterpart).

    FP64 reproduces the reference's all-double march with its operation
    order, every operation separately rounded.  FP64_FAST is the same march
    with fused multiply-adds and reciprocal constants (roundoff-level
    differences).  FP32_BULK keeps the state and runs reconstruction,
    projection, closure, eigenvalues, HLL fluxes and stage algebra in float,
    forms the one cancelling quantity 1 - |J|^2/phi^2 exactly (error-free
    float products and sums), and keeps the step scalars and integrating
    factors, the inflow extrapolation and a compensated dose accumulation in
    double or float pairs.  FP32_BULK_F64_CLOSURE is FP32_BULK with the face
    projection, closure, eigenvalues and HLL combination in double.
    """

    FP64 = auto()
    FP32_BULK = auto()
    FP64_FAST = auto()
    FP32_BULK_F64_CLOSURE = auto()
