"""Beam inflow at the z-min boundary (mirrors ``numerics/boundary.py:11-114``).

The transverse profile is built once on the host, exactly as the reference
builds it, and handed to the device march as ``BeamInflow.spatial``; the
device evaluates the energy-dependent inflow state and the ghost cells itself.
"""

import math
from dataclasses import dataclass
from typing import Any

import numpy as np

from ..beam import Beam
from .grid import Grid


def _gaussian_cell_average(centers, spacing, mean, sigma):
    """Finite-volume cell averages of a unit 1-D Gaussian (boundary.py:11-27)."""
    scale = math.sqrt(2.0) * sigma
    lower = (centers - 0.5 * spacing - mean) / scale
    upper = (centers + 0.5 * spacing - mean) / scale
    mass = np.array(
        [0.5 * (math.erf(b) - math.erf(a)) for a, b in zip(lower, upper)],
        dtype=float,
    )
    return mass / spacing


def _host(a):
    return a.get() if hasattr(a, "get") else np.asarray(a)


def _transverse_profile(beam: Beam, grid: Grid):
    """Unit-mass erf profile with an elliptic cut at cutoff_sigma (boundary.py:30-60)."""
    x = _host(grid.x)
    y = _host(grid.y)
    gx = _gaussian_cell_average(x, grid.dx, beam.center[0], beam.sigma[0])
    gy = _gaussian_cell_average(y, grid.dy, beam.center[1], beam.sigma[1])
    profile = gx[:, None] * gy[None, :]
    r2 = (
        ((x[:, None] - beam.center[0]) / beam.sigma[0]) ** 2
        + ((y[None, :] - beam.center[1]) / beam.sigma[1]) ** 2
    )
    profile = np.where(r2 <= beam.cutoff_sigma**2, profile, 0.0)
    mass = profile.sum() * grid.dx * grid.dy
    if not math.isfinite(mass) or mass <= 0:
        raise ValueError("Beam profile has zero or non-finite mass on the domain.")
    return profile / mass


def energy_density(beam: Beam, energy: float) -> float:
    """Gaussian energy spectrum N(E; E_beam, sigma_E) (boundary.py:85-99)."""
    if beam.energy_spread == 0:
        raise ValueError("Zero energy spread is not supported by the Gaussian inflow.")
    return (
        math.exp(-(energy - beam.energy) ** 2 / (2.0 * beam.energy_spread**2))
        / (math.sqrt(2.0 * math.pi) * beam.energy_spread)
    )


@dataclass(frozen=True)
class BeamInflow:
    beam: Beam
    spatial: Any

    @classmethod
    def from_beam(cls, beam: Beam, grid: Grid, xp):
        spatial = beam.n_protons * _transverse_profile(beam, grid)
        return cls(beam=beam, spatial=xp.asarray(spatial))

    def state(self, energy: float, xp):
        """(phi, 0, 0, exp(-theta^2) phi) at one energy (boundary.py:82-114)."""
        phi = self.spatial * energy_density(self.beam, energy)
        U = xp.empty((4, *phi.shape), dtype=phi.dtype)
        U[0] = phi
        U[1] = 0.0
        U[2] = 0.0
        U[3] = math.exp(-self.beam.angular_spread**2) * phi
        return U
