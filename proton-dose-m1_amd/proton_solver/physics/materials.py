"""Material properties and voxelisation (mirrors ``physics/materials.py:1-137``)."""

from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path
from typing import Any

import numpy as np

from ..enums import Material
from ..geometry import Geometry

DATA_DIR = Path(__file__).resolve().parents[2] / "data"


@lru_cache
def _material_rows():
    """{material value: (density kg/m^3, radiation length m)}."""
    rows = {}
    with open(DATA_DIR / "materials.csv", encoding="utf-8") as fh:
        header = fh.readline().strip().split(",")
        i_mat = header.index("material")
        i_rho = header.index("density_kg_m3")
        i_x0 = header.index("radiation_length_m")
        for line in fh:
            parts = line.strip().split(",")
            if len(parts) < len(header):
                continue
            rows[parts[i_mat]] = (float(parts[i_rho]), float(parts[i_x0]))
    return rows


@dataclass(frozen=True)
class ScaledMaterial:
    """A repository material at `density_scale` times its tabulated density.

    No reference counterpart (the reference's materials are the five tabulated
    ones).  At fixed composition the linear stopping power and the transport
    and removal coefficients are proportional to density, so their tables are
    the base tables times `density_scale`, and the radiation length is divided
    by it.  Used for lung at 0.26 g/cm^3 (the tabulated lung is 1.04).
    """

    base: Material
    density_scale: float

    def __post_init__(self):
        if not isinstance(self.base, Material):
            raise ValueError("base must be a Material.")
        if not (np.isfinite(self.density_scale) and self.density_scale > 0.0):
            raise ValueError("density_scale must be finite and positive.")

    @property
    def value(self) -> str:
        return f"{self.base.value}*{self.density_scale:g}"


def table_material(material):
    """(tabulated material, factor on its linear coefficients)."""
    if isinstance(material, ScaledMaterial):
        return material.base, float(material.density_scale)
    return material, 1.0


def density(material: Material) -> float:
    """Material density in kg/m^3 (materials.py:29-38)."""
    if isinstance(material, ScaledMaterial):
        return density(material.base) * material.density_scale
    row = _material_rows().get(material.value)
    if row is None:
        raise ValueError(f"Unknown material: {material.value}")
    return row[0]


def radiation_length(material: Material) -> float:
    """Radiation length in metres (materials.py:40-63)."""
    if isinstance(material, ScaledMaterial):
        return radiation_length(material.base) / material.density_scale
    row = _material_rows().get(material.value)
    if row is None:
        raise ValueError(f"Unknown material: {material.value}")
    value = row[1]
    if not np.isfinite(value) or value <= 0.0:
        raise ValueError(f"Invalid radiation length for {material.value}: {value}")
    return value


@dataclass(frozen=True)
class MaterialMap:
    ids: Any
    materials: tuple[Material, ...]

    @classmethod
    def from_geometry(cls, geometry: Geometry, grid, xp):
        """Voxelise boxes (half-open on cell centres) or a CT (materials.py:70-123)."""
        if geometry.ct_scan is not None:
            partition = geometry.hu_partition
            materials = tuple(dict.fromkeys(
                (*partition.materials, geometry.background_material)
            ))
            classes = partition.classify(geometry.ct_scan.hu)
            lookup = np.asarray(
                [materials.index(m) for m in partition.materials], dtype=np.int32
            )
            ids = lookup[classes]
            if geometry.ct_scan.body_mask is not None:
                ids = ids.copy()
                ids[~geometry.ct_scan.body_mask] = materials.index(
                    geometry.background_material
                )
            return cls(ids=xp.asarray(ids, dtype=xp.int32), materials=materials)

        materials = tuple(dict.fromkeys(
            [geometry.background_material, *(o.material for o in geometry.objects)]
        ))
        index = {m: i for i, m in enumerate(materials)}
        ids = xp.full(grid.shape, index[geometry.background_material], dtype=xp.int32)
        for box in geometry.objects:
            ix = xp.where((grid.x >= box.x_min) & (grid.x < box.x_max))[0]
            iy = xp.where((grid.y >= box.y_min) & (grid.y < box.y_max))[0]
            iz = xp.where((grid.z >= box.z_min) & (grid.z < box.z_max))[0]
            ids[ix[:, None, None], iy[None, :, None], iz[None, None, :]] = index[box.material]
        return cls(ids=ids, materials=materials)

    def id_for(self, material: Material) -> int:
        return self.materials.index(material)


def density_field(material_map: MaterialMap, xp):
    """Density of every cell (materials.py:130-137)."""
    values = xp.asarray([density(m) for m in material_map.materials])
    return values[material_map.ids]
