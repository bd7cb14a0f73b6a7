"""Cell-centred finite-volume grid (mirrors ``numerics/grid.py:7-61``)."""

from dataclasses import dataclass
from typing import Any

from ..geometry import Geometry


@dataclass(frozen=True)
class Grid:
    x: Any
    y: Any
    z: Any

    dx: float
    dy: float
    dz: float

    shape: tuple[int, int, int]

    @classmethod
    def from_geometry(cls, geometry: Geometry, xp):
        """x and y centred on 0, z from 0 (grid.py:19-42)."""
        nx, ny, nz = geometry.resolution
        lx, ly, lz = geometry.size
        dx, dy, dz = lx / nx, ly / ny, lz / nz
        return cls(
            x=(xp.arange(nx) + 0.5) * dx - 0.5 * lx,
            y=(xp.arange(ny) + 0.5) * dy - 0.5 * ly,
            z=(xp.arange(nz) + 0.5) * dz,
            dx=dx,
            dy=dy,
            dz=dz,
            shape=(nx, ny, nz),
        )

    nx = property(lambda self: self.shape[0])
    ny = property(lambda self: self.shape[1])
    nz = property(lambda self: self.shape[2])

    @property
    def spacing(self):
        return self.dx, self.dy, self.dz

    @property
    def cell_volume(self):
        return self.dx * self.dy * self.dz
