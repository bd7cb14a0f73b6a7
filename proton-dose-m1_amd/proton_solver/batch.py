"""Batched pencil beamlets on a voxel phantom (no reference counterpart).

The reference solves one beam per ``energy_march`` call (solver.py:102-114).
This module marches many beamlets of uneven energy and width through one
phantom in a single device batch: each beamlet is solved by the method of
``energy_march`` on its own box of the phantom (lateral extent fixed, depth
to the beamlet's conservative CSDA range), and the beamlet doses are summed
into the phantom.  The arithmetic is the FP32-bulk march.

Also here: the seeded ellipsoid phantoms and beamlet sets of the batched
workloads, and the problem of one beamlet on its box or on the whole phantom
(for the single-box and full-domain comparisons).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native
from .beam import Beam
from .enums import Material, Precision
from .geometry import Geometry
from .numerics.boundary import BeamInflow
from .numerics.grid import Grid
from .physics.materials import MaterialMap, ScaledMaterial
from .physics.stopping_power import stopping_power_interpolant

# Lung at 0.26 g/cm^3 (the tabulated ICRP lung is 1.04 g/cm^3).
LUNG_026 = ScaledMaterial(Material.LUNG, 0.25)


@dataclass(frozen=True)
class Ellipsoid:
    center: tuple[float, float, float]      # m
    semi_axes: tuple[float, float, float]   # m
    material: object


@dataclass(frozen=True)
class Phantom:
    """Material ids on a cell-centred grid (x, y centred on 0, z from 0)."""

    ids: np.ndarray                          # (nx, ny, nz) int32
    materials: tuple
    spacing: tuple[float, float, float]      # m

    @property
    def shape(self):
        return tuple(int(s) for s in self.ids.shape)

    @property
    def size(self):
        return tuple(n * d for n, d in zip(self.shape, self.spacing))

    def grid(self) -> Grid:
        return Grid.from_geometry(Geometry(size=self.size, resolution=self.shape), np)

    def material_map(self) -> MaterialMap:
        return MaterialMap(ids=self.ids, materials=self.materials)


def ellipsoid_phantom(shape, spacing, ellipsoids, background=Material.WATER) -> Phantom:
    """Voxelise ellipsoids by cell centre; later ellipsoids overwrite earlier ones."""
    materials = tuple(dict.fromkeys([background, *(e.material for e in ellipsoids)]))
    index = {m: i for i, m in enumerate(materials)}
    nx, ny, nz = shape
    dx, dy, dz = spacing
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    y = (np.arange(ny) + 0.5) * dy - 0.5 * ny * dy
    z = (np.arange(nz) + 0.5) * dz
    ids = np.full(shape, index[background], dtype=np.int32)
    for e in ellipsoids:
        (cx, cy, cz), (ax, ay, az) = e.center, e.semi_axes
        r = (((x - cx) / ax) ** 2)[:, None, None] + (((y - cy) / ay) ** 2)[None, :, None] + \
            (((z - cz) / az) ** 2)[None, None, :]
        ids[r <= 1.0] = index[e.material]
    return Phantom(ids=ids, materials=materials, spacing=tuple(float(d) for d in spacing))


def random_ellipsoids(rng, n, size, semi_axis_range=(0.005, 0.030),
                      materials=(Material.BONE_COMPACT, LUNG_026)):
    """n ellipsoids: centres uniform in the phantom, semi-axes uniform in the
    range, material drawn uniformly from `materials`."""
    lx, ly, lz = size
    out = []
    for _ in range(n):
        c = (rng.uniform(-lx / 2, lx / 2), rng.uniform(-ly / 2, ly / 2), rng.uniform(0.0, lz))
        a = tuple(rng.uniform(*semi_axis_range) for _ in range(3))
        m = materials[int(rng.integers(len(materials)))]
        out.append(Ellipsoid(c, a, m))
    return out


@dataclass(frozen=True)
class Beamlet:
    energy: float                       # MeV
    sigma: tuple[float, float]          # m
    center: tuple[float, float]         # m, entry point on the z = 0 face
    energy_spread: float = 1.0          # MeV
    angular_spread: float = 0.0         # rad
    n_protons: float = 1e6
    cutoff_sigma: float = 5.0

    def beam(self, center=None) -> Beam:
        return Beam(n_protons=self.n_protons, energy=self.energy,
                    energy_spread=self.energy_spread, sigma=tuple(self.sigma),
                    center=tuple(self.center if center is None else center),
                    angular_spread=self.angular_spread, cutoff_sigma=self.cutoff_sigma)


def random_beamlets(rng, n, energy_range=(70.0, 230.0), sigma_range=(0.002, 0.006),
                    field=(-0.030, 0.030)):
    """n beamlets: energy, sigma (round) and entry point (x, y) uniform."""
    out = []
    for _ in range(n):
        e = float(rng.uniform(*energy_range))
        s = float(rng.uniform(*sigma_range))
        cx = float(rng.uniform(*field))
        cy = float(rng.uniform(*field))
        out.append(Beamlet(energy=e, sigma=(s, s), center=(cx, cy)))
    return out


def batch_workload(n_beamlets=256, phantom_seed=4, beamlet_seed=40, shape=(200, 200, 300),
                   spacing=(0.001, 0.001, 0.001), n_ellipsoids=12):
    """The 256-beamlet workload: a 200x200x300 water phantom at 1 mm with 12
    seeded ellipsoids (semi-axes 5-30 mm) of cortical bone (1.85 g/cm^3) or
    lung (0.26 g/cm^3), and beamlets with E ~ U(70, 230) MeV, sigma ~ U(2, 6) mm
    and entry points ~ U(-30, 30) mm."""
    size = tuple(n * d for n, d in zip(shape, spacing))
    ell = random_ellipsoids(np.random.default_rng(phantom_seed), n_ellipsoids, size)
    phantom = ellipsoid_phantom(shape, spacing, ell)
    beamlets = random_beamlets(np.random.default_rng(beamlet_seed), n_beamlets)
    return phantom, beamlets


# The five tabulated materials (data/materials.csv), in the order the
# combined workload's ellipsoids draw them.
FIVE_MATERIALS = (Material.AIR, Material.WATER, Material.LUNG, Material.BONE_COMPACT,
                  Material.MUSCLE_SKELETAL)


def combined_workload(n_beamlets=1024, phantom_seed=5, beamlet_seed=50,
                      shape=(256, 256, 200), spacing=(0.001, 0.001, 0.001), n_ellipsoids=40):
    """The combined CT-like workload: a 256x256x200 water phantom at 1 mm with
    40 seeded ellipsoids (semi-axes 5-30 mm) drawn from the five tabulated
    materials (air, water, lung 1.04, cortical bone 1.85, skeletal muscle), and
    1024 beamlets with the 256-beamlet workload's distributions."""
    size = tuple(n * d for n, d in zip(shape, spacing))
    ell = random_ellipsoids(np.random.default_rng(phantom_seed), n_ellipsoids, size,
                            materials=FIVE_MATERIALS)
    phantom = ellipsoid_phantom(shape, spacing, ell)
    beamlets = random_beamlets(np.random.default_rng(beamlet_seed), n_beamlets)
    return phantom, beamlets


def air_free_beamlets(phantom: Phantom, beamlets, boxes, margin_cells=48):
    """Indices of the beamlets with no air within `margin_cells` cells of the
    entry column (laterally) over their box depth: under the vacuum-aware CFL
    their energy steps are not bound by air (whose step is 1/934 of water's)."""
    if Material.AIR not in phantom.materials:
        return list(range(len(beamlets)))
    air = phantom.ids == phantom.materials.index(Material.AIR)
    px, py, _ = phantom.shape
    dx, dy, _ = phantom.spacing
    out = []
    for b, (bl, box) in enumerate(zip(beamlets, boxes)):
        ix = int(np.floor((bl.center[0] + 0.5 * px * dx) / dx))
        iy = int(np.floor((bl.center[1] + 0.5 * py * dy) / dy))
        nz = box[4]
        if not air[max(ix - margin_cells, 0):ix + margin_cells,
                   max(iy - margin_cells, 0):iy + margin_cells, :nz].any():
            out.append(b)
    return out


# Tile threshold of the batched workloads, relative to a beamlet's peak inflow
# fluence: tiles whose 9-cell neighbourhood stays below it are left out and
# held at exact vacuum.  1e-6 lies below the truncation of the reference's
# moving active domain (exp(-5^2/2) = 3.7e-6 of the current maximum,
# energy.py:1301-1303); against plain FP64 marches it leaves the maximum-dose
# error (<= 6.4e-4) and gamma(1 %/1 mm) (100 %) of the 256-beamlet
# benchmark's stress beamlets where 1e-10 had them, with 21 % fewer tiles.
BATCH_TILE_THRESHOLD = 1e-6


def batch_workload_config(tile_threshold=BATCH_TILE_THRESHOLD, vacuum_aware_cfl=False,
                          precision=Precision.FP32_BULK):
    """Physics and numerics of the batched workloads: Wentzel scattering and
    nuclear removal, WENO-Z, IF-SSPRK3 with exact M1 speeds, max step
    0.25 MeV, CFL 0.9 over the whole phantom (as the march of one beamlet on
    the phantom takes it; vacuum_aware_cfl selects the reference's
    vacuum-aware rule, energy.py:459-537, instead), energy_min 1 MeV,
    FP32-bulk (or `precision`) with tile skipping at BATCH_TILE_THRESHOLD.  Its geometry and
    beam are placeholders: the batch takes the phantom and the beamlets."""
    from .beam import Beam as _Beam
    from .config import DeviceOptions, Numerics, Physics, SolverConfig
    from .enums import EnergyScheme, Scattering, SpatialScheme, WaveSpeeds

    return SolverConfig(
        geometry=Geometry(size=(0.2, 0.2, 0.3), resolution=(200, 200, 300)),
        beam=_Beam(energy=150.0),
        physics=Physics(scattering=Scattering.WENTZEL, nuclear=True),
        numerics=Numerics(spatial_scheme=SpatialScheme.WENO_Z,
                          energy_scheme=EnergyScheme.IF_SSPRK3, wave_speeds=WaveSpeeds.EXACT,
                          max_energy_step=0.25, cfl=0.9, energy_min=1.0,
                          vacuum_aware_cfl=vacuum_aware_cfl),
        device=DeviceOptions(precision=precision, tile_threshold=tile_threshold),
    )


# ------------------------------------------------------------------ boxes
def _residual_range_tables(materials, e_lo, e_hi, n=4000):
    """Per material: energies and CSDA residual range R(E) = int_{e_lo}^E dE'/S."""
    from .physics.coefficient_interpolation import pchip_evaluate

    grid = np.linspace(e_lo, e_hi, n)
    tables = []
    for m in materials:
        E, v, s = stopping_power_interpolant(m)
        S = np.array([pchip_evaluate(E, v, s, float(e)) for e in grid])
        inv = 1.0 / S
        R = np.concatenate([[0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(grid))])
        tables.append(R)
    return grid, tables


def box_depths(phantom: Phantom, boxes_xy, energies, energy_min, guard_cells=10):
    """Depth (cells) to the conservative CSDA range of each beamlet in its box.

    Each z slice is crossed with the least-stopping material present in the
    box's cross-section there, so no path through the box goes deeper; the
    guard adds cells for the distal tail of the march."""
    nx, ny, nz = phantom.shape
    dz = phantom.spacing[2]
    mats = phantom.materials
    e_hi = float(max(energies)) + 1.0
    grid, R = _residual_range_tables(mats, max(energy_min, 0.5) * 0.5, e_hi)
    B = len(boxes_xy)
    present = np.zeros((B, len(mats), nz), dtype=bool)
    for b, (x0, y0, bx, by) in enumerate(boxes_xy):
        sub = phantom.ids[x0:x0 + bx, y0:y0 + by, :]
        for m in range(len(mats)):
            present[b, m] = (sub == m).any(axis=(0, 1))
    E = np.asarray(energies, dtype=np.float64).copy()
    depth = np.full(B, nz, dtype=np.int64)
    alive = np.ones(B, dtype=bool)
    for z in range(nz):
        best = np.full(B, -np.inf)
        for m in range(len(mats)):
            r = np.interp(E, grid, R[m]) - dz
            e_next = np.where(r > 0.0, np.interp(np.maximum(r, 0.0), R[m], grid), -np.inf)
            best = np.where(present[:, m, z], np.maximum(best, e_next), best)
        stop = alive & (best <= energy_min)
        depth[stop] = z + 1
        alive &= ~stop
        E = np.where(alive, best, E)
        if not alive.any():
            break
    return np.minimum(depth + guard_cells, nz)


def beamlet_boxes(phantom: Phantom, beamlets, energy_min=1.0, window_sigma=5.0,
                  lateral=(64, 64), guard_cells=10, z_multiple=32):
    """(x0, y0, nx, ny, nz) of each beamlet's box: `lateral` cells centred on
    the cell holding the entry point (clamped into the phantom), depth to the
    conservative range of E + window_sigma * sigma_E, rounded up to a multiple
    of `z_multiple` cells (a multiple of the workgroup tile depth) within the
    phantom."""
    px, py, pz = phantom.shape
    dx, dy, _ = phantom.spacing
    bx, by = (min(int(v), p) for v, p in zip(lateral, (px, py)))
    xy = []
    for bl in beamlets:
        ix = int(np.floor((bl.center[0] + 0.5 * px * dx) / dx))
        iy = int(np.floor((bl.center[1] + 0.5 * py * dy) / dy))
        x0 = min(max(ix - bx // 2, 0), px - bx)
        y0 = min(max(iy - by // 2, 0), py - by)
        xy.append((x0, y0, bx, by))
    e0 = [bl.energy + window_sigma * bl.energy_spread for bl in beamlets]
    depth = box_depths(phantom, xy, e0, energy_min, guard_cells)
    boxes = []
    for (x0, y0, bx_, by_), d in zip(xy, depth):
        nz = int(min(pz, -(-int(d) // z_multiple) * z_multiple))
        boxes.append((x0, y0, bx_, by_, max(nz, 4)))
    return boxes


def box_problem(phantom: Phantom, box, beamlet: Beamlet):
    """(grid, material map, inflow, beam) of one beamlet on its box alone."""
    x0, y0, nx, ny, nz = box
    px, py, pz = phantom.shape
    if min(nx, ny, nz) <= 0 or x0 < 0 or y0 < 0 or x0 + nx > px or y0 + ny > py or nz > pz:
        raise ValueError(f"box {tuple(box)} lies outside the phantom")
    dx, dy, dz = phantom.spacing
    grid = Grid.from_geometry(Geometry(size=(nx * dx, ny * dy, nz * dz), resolution=(nx, ny, nz)),
                              np)
    # box centre in phantom coordinates
    cx = (x0 + 0.5 * nx) * dx - 0.5 * px * dx
    cy = (y0 + 0.5 * ny) * dy - 0.5 * py * dy
    beam = beamlet.beam(center=(beamlet.center[0] - cx, beamlet.center[1] - cy))
    mm = MaterialMap(ids=np.ascontiguousarray(phantom.ids[x0:x0 + nx, y0:y0 + ny, :nz]),
                     materials=phantom.materials)
    return grid, mm, BeamInflow.from_beam(beam, grid, np), beam


def full_problem(phantom: Phantom, beamlet: Beamlet):
    """(grid, material map, inflow, beam) of one beamlet on the whole phantom."""
    grid = phantom.grid()
    beam = beamlet.beam()
    return grid, phantom.material_map(), BeamInflow.from_beam(beam, grid, np), beam


# ------------------------------------------------------------------ march
@dataclass
class BatchResult:
    dose: np.ndarray | None        # phantom dose, Gy (sum over beamlets)
    steps: np.ndarray              # energy steps per beamlet
    boxes: list
    handle: _native.Batch = field(repr=False, default=None)

    def beamlet_dose(self, b: int) -> np.ndarray:
        """Dose of beamlet b on its box, (nx_b, ny_b, nz_b)."""
        return self.handle.beamlet_dose(b)


def prepare_batch(phantom: Phantom, beamlets, config, boxes=None, device=None, tile_shape=0):
    """Bind a phantom and its beamlets to a device batch (returns it with the
    boxes); tile_shape indexes _native.batch_tile_shapes()."""
    n = config.numerics
    if boxes is None:
        boxes = beamlet_boxes(phantom, beamlets, n.energy_min, n.energy_window_sigma)
    dev = config.device.device if device is None else device
    handle = _native.Batch(dev)
    if tile_shape:
        handle.set_tile_shape(tile_shape)
    profiles, beams = [], []
    for box, bl in zip(boxes, beamlets):
        _, _, inflow, beam = box_problem(phantom, box, bl)
        profiles.append(np.asarray(inflow.spatial))
        beams.append(beam)
    handle.set_problem(phantom.ids, phantom.spacing, phantom.materials, boxes, profiles, beams)
    return handle, boxes


def batch_march(phantom: Phantom, beamlets, config, boxes=None, want_dose=True,
                handle=None) -> BatchResult:
    """March every beamlet through the phantom; the dose is their sum.

    `config` supplies physics and numerics (a SolverConfig; its geometry and
    beam are not used).  The batch runs the FP32-bulk march."""
    if config.device.precision not in (Precision.FP32_BULK, Precision.FP32_BULK_F64_CLOSURE):
        raise NotImplementedError("the batched march runs the FP32-bulk kernels "
                                  "(DeviceOptions(precision=Precision.FP32_BULK))")
    if handle is None:
        handle, boxes = prepare_batch(phantom, beamlets, config, boxes)
    num = _native.numerics_struct(config)
    dose, steps = handle.march(num, want_dose)
    return BatchResult(dose=dose, steps=steps, boxes=list(handle.boxes), handle=handle)


# ------------------------------------------------------- one beam, whole grid
def interface_workload(offset_m=0.0, shape=(160, 160, 480), spacing=(0.0005,) * 3):
    """The lateral water / cortical-bone interface workload: a 160 x 160 x 480
    grid at 0.5 mm whose x > 0 half is bone (1.85 g/cm^3), and one 200 MeV
    beam (sigma 3 mm, sigma_E 1 MeV) entering at x = -2 mm + offset_m, y = 0.
    Returned as (phantom, [beamlet], [box]) with the box the whole grid, so
    the batched march (and its tile-skip network) solves it as one beamlet
    on the full domain."""
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) * spacing[0] - 0.5 * nx * spacing[0]
    ids = np.zeros(shape, dtype=np.int32)
    ids[x >= 0.0] = 1  # half-open on cell centres, as a Box(0, ...) voxelises
    phantom = Phantom(ids=ids, materials=(Material.WATER, Material.BONE_COMPACT),
                      spacing=tuple(float(d) for d in spacing))
    beamlet = Beamlet(energy=200.0, sigma=(0.003, 0.003), center=(-0.002 + offset_m, 0.0))
    return phantom, [beamlet], [(0, 0, nx, ny, nz)]


def interface_offsets(n, seed=3, spread_m=0.004):
    """Seeded beam offsets along x, U(-spread, spread) (the training runs of
    the interface workload's tile-skip network)."""
    rng = np.random.default_rng(seed)
    return [float(v) for v in rng.uniform(-spread_m, spread_m, n)]
