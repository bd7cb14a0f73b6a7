"""Angular scattering coefficients (mirrors ``physics/scattering.py:14-342``)."""

from functools import lru_cache
from pathlib import Path

import numpy as np

from ..enums import CoefficientInterpolation, Material, Scattering
from .coefficient_interpolation import check_in_table, interpolate_coefficient, pchip_slopes
from .materials import MaterialMap, ScaledMaterial, radiation_length
from .stopping_power import _read_table, stopping_power

DATA_DIR = Path(__file__).resolve().parents[2] / "data" / "wentzel_transport"

PROTON_REST_ENERGY_MEV = 938.27208816
ROSSI_ES_MEV = 15.0


@lru_cache
def load_transport_table(material: Material):
    """(energies [MeV], Sigma_tr [1/m]) of the Wentzel-VI table."""
    if isinstance(material, ScaledMaterial):
        energies, values = load_transport_table(material.base)
        return energies, values * material.density_scale
    return _read_table(
        DATA_DIR / f"{material.value}.csv",
        ("energy_mev", "first_transport_coefficient_per_m"),
    )


@lru_cache
def transport_table_slopes(material: Material):
    energies, coefficients = load_transport_table(material)
    return pchip_slopes(energies, coefficients)


def first_transport_coefficient(
    material: Material,
    energy: float,
    interpolation: CoefficientInterpolation = CoefficientInterpolation.LINEAR,
) -> float:
    """Wentzel first transport coefficient [1/m] (scattering.py:65-116)."""
    energies, coefficients = load_transport_table(material)
    energy = check_in_table(energy, energies, f"the Wentzel table for {material.value}")
    return interpolate_coefficient(
        energies, coefficients, transport_table_slopes(material), energy, interpolation
    )


def proton_pv_mev(energy: float) -> float:
    """Relativistic p*v of a proton in MeV (scattering.py:118-134)."""
    energy = max(float(energy), 1.0e-8)
    m = PROTON_REST_ENERGY_MEV
    return energy * (energy + 2.0 * m) / (energy + m)


def rossi_transport_coefficient(material: Material, energy: float) -> float:
    """(Es / pv)^2 / X0 (scattering.py:137-153)."""
    return (ROSSI_ES_MEV / proton_pv_mev(energy)) ** 2 / radiation_length(material)


def transport_coefficient(material, energy, model, interpolation=CoefficientInterpolation.LINEAR):
    if model is Scattering.WENTZEL:
        return first_transport_coefficient(material, energy, interpolation)
    if model is Scattering.ROSSI:
        return rossi_transport_coefficient(material, energy)
    if model is Scattering.OFF:
        return 0.0
    raise ValueError(f"Unknown scattering model: {model}")


def scattering_rate_field(material_map: MaterialMap, energy, model, xp,
                          interpolation=CoefficientInterpolation.LINEAR):
    """Sigma_tr(E) / S(E), scalar for one material (scattering.py:227-265)."""
    rates = [
        transport_coefficient(m, energy, model, interpolation) / stopping_power(m, energy)
        for m in material_map.materials
    ]
    if len(rates) == 1:
        return rates[0]
    return xp.asarray(rates)[material_map.ids]


def scattering_optical_depth(material_map: MaterialMap, energy, dE, model, xp,
                             interpolation=CoefficientInterpolation.LINEAR):
    """Two-point Gauss-Legendre integral of Sigma_tr/S (scattering.py:267-306)."""
    if model is Scattering.OFF:
        return xp.zeros(material_map.ids.shape, dtype=float)
    midpoint = energy - 0.5 * dE
    half_width = 0.5 * dE
    offset = half_width / np.sqrt(3.0)
    return half_width * (
        scattering_rate_field(material_map, midpoint - offset, model, xp, interpolation)
        + scattering_rate_field(material_map, midpoint + offset, model, xp, interpolation)
    )
