"""Electronic stopping power S(E) (mirrors ``physics/stopping_power.py:10-242``)."""

from functools import lru_cache
from pathlib import Path

import numpy as np

from ..enums import Material
from .coefficient_interpolation import check_in_table, pchip_evaluate, pchip_slopes
from .materials import MaterialMap, ScaledMaterial

DATA_DIR = Path(__file__).resolve().parents[2] / "data" / "stopping_power"


def _read_table(path: Path, columns: tuple[str, ...]):
    """Columns of a CSV table with a header line, as float64 arrays."""
    with open(path, encoding="utf-8") as fh:
        header = fh.readline().strip().split(",")
    idx = [header.index(c) for c in columns]
    data = np.loadtxt(path, delimiter=",", skiprows=1, dtype=float, ndmin=2)
    return tuple(np.ascontiguousarray(data[:, i]) for i in idx)


@lru_cache
def load_stopping_power_table(material: Material):
    """(energies [MeV], S [MeV/m]) of one material."""
    if isinstance(material, ScaledMaterial):
        energies, values = load_stopping_power_table(material.base)
        return energies, values * material.density_scale
    path = DATA_DIR / f"{material.value}.csv"
    if not path.exists():
        raise FileNotFoundError(f"No stopping-power data found for {material.value}.")
    return _read_table(path, ("energy_mev", "electronic_stopping_power_mev_per_m"))


@lru_cache
def stopping_power_interpolant(material: Material):
    """(energies, values, PCHIP slopes) after the checks of stopping_power.py:73-91."""
    energies, values = load_stopping_power_table(material)
    if len(energies) < 2:
        raise ValueError(
            f"Stopping-power table for {material.value} must contain at least two points."
        )
    if not np.all(np.diff(energies) > 0.0):
        raise ValueError(
            f"Stopping-power energies for {material.value} must be strictly increasing."
        )
    if not np.all(np.isfinite(values)) or not np.all(values > 0.0):
        raise ValueError(
            f"Stopping-power values for {material.value} must be finite and positive."
        )
    return energies, values, pchip_slopes(energies, values)


def stopping_power(material: Material, energy: float) -> float:
    """PCHIP-interpolated stopping power at one energy (stopping_power.py:141-228)."""
    energies, values, slopes = stopping_power_interpolant(material)
    energy = check_in_table(energy, energies, "stopping-power table")
    return pchip_evaluate(energies, values, slopes, energy)


def stopping_power_field(material_map: MaterialMap, energy: float, xp):
    values = xp.asarray([stopping_power(m, energy) for m in material_map.materials])
    return values[material_map.ids]
