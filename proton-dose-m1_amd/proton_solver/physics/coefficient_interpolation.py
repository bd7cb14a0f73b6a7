"""Monotone cubic (PCHIP) and linear interpolation of tabulated coefficients.

Mirrors ``physics/coefficient_interpolation.py:19-92`` and the stopping-power
interpolant of ``physics/stopping_power.py:30-138``: Fritsch-Carlson weighted
harmonic-mean interior slopes, the SciPy-style shape-preserving end-point rule,
and cubic Hermite evaluation on the interval found by a right-sided search.
"""

from __future__ import annotations

import numpy as np

from ..enums import CoefficientInterpolation


def _endpoint_slope(h0: float, h1: float, d0: float, d1: float) -> float:
    """Shape-preserving one-sided slope (stopping_power.py:30-54)."""
    slope = ((2.0 * h0 + h1) * d0 - h0 * d1) / (h0 + h1)
    if np.sign(slope) != np.sign(d0):
        return 0.0
    if np.sign(d0) != np.sign(d1) and abs(slope) > 3.0 * abs(d0):
        return 3.0 * d0
    return slope


def pchip_slopes(energies, values):
    """Nodal PCHIP slopes of a table."""
    e = np.asarray(energies, dtype=float)
    v = np.asarray(values, dtype=float)
    h = np.diff(e)
    d = np.diff(v) / h
    slopes = np.zeros_like(v)
    if len(e) == 2:
        slopes[:] = d[0]
        return slopes
    # Interior nodes whose neighbouring secants share a sign get the weighted
    # harmonic mean; the others keep slope 0.
    keep = d[:-1] * d[1:] > 0.0
    w_left = 2.0 * h[1:] + h[:-1]
    w_right = h[1:] + 2.0 * h[:-1]
    inner = slopes[1:-1]
    inner[keep] = (w_left[keep] + w_right[keep]) / (
        w_left[keep] / d[:-1][keep] + w_right[keep] / d[1:][keep]
    )
    slopes[0] = _endpoint_slope(h[0], h[1], d[0], d[1])
    slopes[-1] = _endpoint_slope(h[-1], h[-2], d[-1], d[-2])
    return slopes


def pchip_evaluate(energies, values, slopes, energy: float) -> float:
    """Cubic Hermite value at one energy already inside the table."""
    if energy >= energies[-1]:
        return float(values[-1])
    i = int(np.searchsorted(energies, energy, side="right") - 1)
    i = max(0, min(i, len(energies) - 2))
    h = energies[i + 1] - energies[i]
    t = (energy - energies[i]) / h
    t2 = t * t
    t3 = t2 * t
    return float(
        (2.0 * t3 - 3.0 * t2 + 1.0) * values[i]
        + (t3 - 2.0 * t2 + t) * h * slopes[i]
        + (-2.0 * t3 + 3.0 * t2) * values[i + 1]
        + (t3 - t2) * h * slopes[i + 1]
    )


def interpolate_coefficient(energies, values, slopes, energy: float,
                            interpolation: CoefficientInterpolation) -> float:
    if interpolation is CoefficientInterpolation.LINEAR:
        return float(np.interp(energy, energies, values))
    if interpolation is CoefficientInterpolation.PCHIP:
        return pchip_evaluate(energies, values, slopes, energy)
    raise ValueError(f"Unknown coefficient interpolation: {interpolation}")


def check_in_table(energy: float, energies, what: str) -> float:
    """Reject energies outside the table (up to np.isclose) and clamp."""
    if (energy < energies[0] and not np.isclose(energy, energies[0])) or (
        energy > energies[-1] and not np.isclose(energy, energies[-1])
    ):
        raise ValueError(
            f"Energy {energy} MeV outside {what} range "
            f"[{energies[0]}, {energies[-1]}] MeV."
        )
    return float(np.clip(energy, energies[0], energies[-1]))
