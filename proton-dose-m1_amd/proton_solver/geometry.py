"""Domain geometry: the ``proton_solver.geometry`` interface
(reference ``src/proton_solver/geometry.py:10-271``).

The public classes keep the reference's fields, defaults, half-open
conventions, exception types and messages, since callers construct them
directly.  The implementation is this package's own: validation goes
through a few shared predicates, a CT file is read by ``read_ct_npz`` and
oriented by ``orient_volume``, and box overlaps are found with a sweep over
the boxes sorted along x.
"""

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .enums import Material

# Array names a CT NPZ may use for its HU volume, in order of preference; a
# file holding one unnamed array (np.savez(path, array)) stores it as arr_0.
_HU_KEYS = ("HU", "hu")
_UNNAMED_KEY = "arr_0"


def _require(ok: bool, message: str):
    if not ok:
        raise ValueError(message)


def _all_finite(values) -> bool:
    return all(math.isfinite(v) for v in values)


def _positive_finite_triple(values) -> bool:
    return len(values) == 3 and all(math.isfinite(v) and v > 0 for v in values)


def _check_hu_volume(hu: np.ndarray):
    _require(hu.ndim == 3 and np.issubdtype(hu.dtype, np.number),
             "CT HU must be a numeric 3D array.")
    _require(bool(np.isfinite(hu).all()), "CT HU contains non-finite values.")


@dataclass(frozen=True)
class Box:
    """Axis-aligned box using half-open intervals [min, max), in metres."""

    x_min: float
    x_max: float
    y_min: float
    y_max: float
    z_min: float
    z_max: float

    material: Material

    def __post_init__(self):
        _require(_all_finite(self.bounds()), "Box coordinates must be finite.")
        _require(all(lo < hi for lo, hi in self.intervals()),
                 "Box minima must be smaller than maxima.")

    def bounds(self):
        return (self.x_min, self.x_max, self.y_min, self.y_max, self.z_min, self.z_max)

    def intervals(self):
        """((x_min, x_max), (y_min, y_max), (z_min, z_max))."""
        b = self.bounds()
        return ((b[0], b[1]), (b[2], b[3]), (b[4], b[5]))

    def overlaps(self, other: "Box") -> bool:
        """Two half-open boxes share a point when they overlap on every axis."""
        return all(a_lo < b_hi and b_lo < a_hi
                   for (a_lo, a_hi), (b_lo, b_hi) in zip(self.intervals(), other.intervals()))


@dataclass(frozen=True)
class HUPartition:
    """Half-open HU bins mapped to transport materials: bin i holds
    cutoffs[i-1] <= HU < cutoffs[i] and is transported as materials[i]."""

    cutoffs: tuple[float, ...] = (-200.0, 150.0)
    materials: tuple[Material, ...] = (
        Material.LUNG,
        Material.WATER,
        Material.BONE_COMPACT,
    )

    def __post_init__(self):
        _require(len(self.materials) == len(self.cutoffs) + 1,
                 "HUPartition needs one more material than cutoffs.")
        _require(all(isinstance(m, Material) for m in self.materials),
                 "HUPartition materials must be Material values.")
        _require(_all_finite(self.cutoffs), "HU cutoffs must be finite.")
        _require(all(np.diff(np.asarray(self.cutoffs, dtype=float)) > 0),
                 "HU cutoffs must be strictly increasing.")

    def classify(self, hu: np.ndarray) -> np.ndarray:
        """Bin index of every voxel (int32): the number of cutoffs <= HU."""
        hu = np.asarray(hu)
        _check_hu_volume(hu)
        return np.searchsorted(self.cutoffs, hu, side="right").astype(np.int32)


def read_ct_npz(path):
    """(HU volume, body mask or None) stored in a CT NPZ, unpickled never."""
    with np.load(path, allow_pickle=False) as archive:
        names = archive.files
        key = next((k for k in _HU_KEYS if k in names), None)
        if key is None and names == [_UNNAMED_KEY]:
            key = _UNNAMED_KEY
        _require(key is not None, "CT NPZ needs a 'HU', 'hu', or single 'arr_0' array.")
        hu = np.asarray(archive[key])
        mask = np.asarray(archive["body_mask"]) if "body_mask" in names else None
    return hu, mask


def orient_volume(volume: np.ndarray, axis_order, flip_axes) -> np.ndarray:
    """Solver-oriented copy of a volume: its axes permuted to axis_order,
    then reversed along every solver axis flagged in flip_axes."""
    out = np.transpose(volume, tuple(axis_order))
    flipped = tuple(a for a, f in enumerate(flip_axes) if f)
    if flipped:
        out = np.flip(out, axis=flipped)
    return np.ascontiguousarray(out)


@dataclass(frozen=True)
class CTScan:
    """A CT volume read from an NPZ (``HU``, ``hu`` or a single ``arr_0``
    array, optional ``body_mask``) and oriented as solver (x, y, +z).
    source_spacing_m gives the voxel size of the stored axes, since the file
    need not hold it."""

    path: str | Path
    source_spacing_m: tuple[float, float, float]
    axis_order: tuple[int, int, int] = (0, 1, 2)
    flip_axes: tuple[bool, bool, bool] = (False, False, False)

    hu: np.ndarray = field(init=False, repr=False, compare=False)
    body_mask: np.ndarray | None = field(init=False, repr=False, compare=False)
    spacing: tuple[float, float, float] = field(init=False)

    def __post_init__(self):
        _require(_positive_finite_triple(self.source_spacing_m),
                 "CT spacing must contain three positive finite metres.")
        _require(sorted(self.axis_order) == [0, 1, 2],
                 "axis_order must be a permutation of (0, 1, 2).")
        _require(len(self.flip_axes) == 3 and all(type(f) is bool for f in self.flip_axes),
                 "flip_axes must contain three booleans.")

        hu, mask = read_ct_npz(self.path)
        _check_hu_volume(hu)
        _require(mask is None or mask.shape == hu.shape,
                 "body_mask shape must match the HU array.")

        oriented = {
            "hu": orient_volume(hu, self.axis_order, self.flip_axes),
            "body_mask": (None if mask is None else
                          orient_volume(mask.astype(bool), self.axis_order, self.flip_axes)),
            "spacing": tuple(self.source_spacing_m[a] for a in self.axis_order),
        }
        for name, value in oriented.items():
            object.__setattr__(self, name, value)

    @property
    def shape(self) -> tuple[int, int, int]:
        return self.hu.shape


def _first_overlap(boxes):
    """A pair of overlapping boxes, or None: boxes sorted by x_min are swept
    along x, each compared only with the earlier boxes whose x interval is
    still open."""
    order = sorted(range(len(boxes)), key=lambda i: boxes[i].x_min)
    open_boxes = []
    for i in order:
        b = boxes[i]
        open_boxes = [a for a in open_boxes if a.x_max > b.x_min]
        for a in open_boxes:
            if a.overlaps(b):
                return a, b
        open_boxes.append(b)
    return None


@dataclass
class Geometry:
    size: tuple[float, float, float]                # m
    resolution: tuple[int, int, int]

    background_material: Material = Material.WATER
    objects: list[Box] = field(default_factory=list)

    view_geometry: bool = False

    ct_scan: CTScan | None = None
    hu_partition: HUPartition = field(default_factory=HUPartition)

    def __post_init__(self):
        _require(len(self.size) == 3 and len(self.resolution) == 3,
                 "size and resolution must contain exactly three values.")
        _require(_positive_finite_triple(self.size),
                 "Geometry dimensions must be finite and positive.")
        _require(all(type(n) is int and n > 0 for n in self.resolution),
                 "Geometry resolution must contain positive integers.")
        self._verify_objects()
        if self.ct_scan is not None:
            self._verify_ct()

    def _verify_objects(self):
        """Every box inside the domain (x, y centred on 0, z from 0), no two
        overlapping."""
        half_x, half_y = self.size[0] / 2, self.size[1] / 2
        domain = ((-half_x, half_x), (-half_y, half_y), (0.0, self.size[2]))
        for box in self.objects:
            inside = all(lo <= b_lo < b_hi <= hi
                         for (b_lo, b_hi), (lo, hi) in zip(box.intervals(), domain))
            _require(inside, "Geometry object extends outside the domain.")
        _require(_first_overlap(self.objects) is None, "Geometry objects may not overlap.")

    def _verify_ct(self):
        scan = self.ct_scan
        _require(not self.objects, "CT and Box objects cannot be combined.")
        _require(self.resolution == scan.shape,
                 "Geometry resolution must match the CT array shape.")
        _require(all(math.isclose(a, b, rel_tol=1e-9, abs_tol=1e-12)
                     for a, b in zip(self.spacing, scan.spacing)),
                 "Geometry cell spacing must match the CT spacing.")
        _require(not self.view_geometry,
                 "3D Box viewer does not support CT; set view_geometry=False.")

    @classmethod
    def from_ct_scan(
        cls,
        path,
        *,
        source_spacing_m,
        axis_order=(0, 1, 2),
        flip_axes=(False, False, False),
        hu_partition=None,
        background_material=Material.WATER,
    ) -> "Geometry":
        """The domain of an oriented CT volume: one cell per voxel."""
        scan = CTScan(path=path, source_spacing_m=source_spacing_m, axis_order=axis_order,
                      flip_axes=flip_axes)
        extent = tuple(n * d for n, d in zip(scan.shape, scan.spacing))
        return cls(size=extent, resolution=scan.shape, background_material=background_material,
                   ct_scan=scan,
                   hu_partition=HUPartition() if hu_partition is None else hu_partition)

    @staticmethod
    def _boxes_overlap(a: Box, b: Box) -> bool:
        return a.overlaps(b)

    @property
    def nx(self):
        return self.resolution[0]

    @property
    def ny(self):
        return self.resolution[1]

    @property
    def nz(self):
        return self.resolution[2]

    @property
    def lx(self):
        return self.size[0]

    @property
    def ly(self):
        return self.size[1]

    @property
    def lz(self):
        return self.size[2]

    @property
    def dx(self):
        return self.lx / self.nx

    @property
    def dy(self):
        return self.ly / self.ny

    @property
    def dz(self):
        return self.lz / self.nz

    @property
    def spacing(self):
        return self.dx, self.dy, self.dz

    @property
    def shape(self):
        return self.resolution
