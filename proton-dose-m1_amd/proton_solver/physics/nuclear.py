"""Primary-proton nuclear removal (mirrors ``physics/nuclear.py:23-252``)."""

from __future__ import annotations

from functools import lru_cache
from pathlib import Path

import numpy as np

from ..enums import CoefficientInterpolation, Material
from .coefficient_interpolation import check_in_table, interpolate_coefficient, pchip_slopes
from .materials import MaterialMap, ScaledMaterial
from .stopping_power import _read_table, stopping_power

DATA_DIR = Path(__file__).resolve().parents[2] / "data" / "nuclear_attenuation"


@lru_cache
def load_nuclear_table(material: Material):
    """(energies [MeV], Sigma_r [1/m]) with the checks of nuclear.py:30-81."""
    if isinstance(material, ScaledMaterial):
        energies, values = load_nuclear_table(material.base)
        return energies, values * material.density_scale
    path = DATA_DIR / f"{material.value}.csv"
    if not path.exists():
        raise FileNotFoundError(f"No nuclear-attenuation data found for {material.value}.")
    energies, coefficients = _read_table(
        path, ("energy_mev", "nuclear_removal_coefficient_per_m")
    )
    if len(energies) < 2:
        raise ValueError(
            f"Nuclear-attenuation table for {material.value} must contain at least two points."
        )
    if not np.all(np.diff(energies) > 0.0):
        raise ValueError(
            f"Nuclear-attenuation energies for {material.value} must be strictly increasing."
        )
    if not np.all(np.isfinite(coefficients)) or np.any(coefficients < 0.0):
        raise ValueError(
            f"Nuclear-removal coefficients for {material.value} must be finite and non-negative."
        )
    return energies, coefficients


@lru_cache
def nuclear_table_slopes(material: Material):
    energies, coefficients = load_nuclear_table(material)
    return pchip_slopes(energies, coefficients)


def nuclear_removal_coefficient(
    material: Material,
    energy: float,
    interpolation: CoefficientInterpolation = CoefficientInterpolation.LINEAR,
) -> float:
    """Proton-inelastic macroscopic removal coefficient [1/m] (nuclear.py:94-138)."""
    energies, coefficients = load_nuclear_table(material)
    energy = check_in_table(energy, energies, f"the nuclear table for {material.value}")
    return interpolate_coefficient(
        energies, coefficients, nuclear_table_slopes(material), energy, interpolation
    )


def nuclear_attenuation_rate_field(material_map: MaterialMap, energy, xp,
                                   interpolation=CoefficientInterpolation.LINEAR):
    rates = [
        nuclear_removal_coefficient(m, energy, interpolation) / stopping_power(m, energy)
        for m in material_map.materials
    ]
    if len(rates) == 1:
        return rates[0]
    return xp.asarray(rates)[material_map.ids]


def nuclear_optical_depth(material_map: MaterialMap, energy, dE, xp,
                          interpolation=CoefficientInterpolation.LINEAR):
    """Two-point Gauss-Legendre integral of Sigma_r/S (nuclear.py:202-233)."""
    if dE <= 0.0:
        raise ValueError("dE must be positive.")
    midpoint = energy - 0.5 * dE
    half_width = 0.5 * dE
    offset = half_width / np.sqrt(3.0)
    return half_width * (
        nuclear_attenuation_rate_field(material_map, midpoint - offset, xp, interpolation)
        + nuclear_attenuation_rate_field(material_map, midpoint + offset, xp, interpolation)
    )
