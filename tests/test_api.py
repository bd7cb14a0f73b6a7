"""Host-side API of the drop-in package (CPU only, no device calls).

Configuration validation, geometry voxelisation, inflow profile, physics
scalars and the dose metrics keep the reference's behaviour
(config.py:53-209, geometry.py:23-271, materials.py:65-137,
stopping_power.py:141-228, metrics.py:146-460).
"""

import math

import numpy as np
import pytest

from proton_solver import (
    Backend, Beam, Box, DeviceOptions, EnergyScheme, Geometry, HUPartition, Material,
    Numerics, Physics, Precision, Scattering, SolverConfig,
)
from proton_solver import metrics
from proton_solver.numerics.grid import Grid
from proton_solver.physics.materials import MaterialMap, density, density_field
from proton_solver.physics.stopping_power import stopping_power


def geo(**kw):
    return Geometry(size=(0.02, 0.02, 0.05), resolution=(10, 10, 25), **kw)


@pytest.mark.parametrize("kwargs, msg", [
    (dict(max_energy_step=0.0), "max_energy_step must be positive"),
    (dict(cfl=1.5), "cfl must lie in"),
    (dict(energy_min=-1.0), "energy_min cannot be negative"),
    (dict(energy_window_sigma=0.0), "energy_window_sigma must be positive"),
    (dict(active_guard_cells=3), "active_guard_cells"),
    (dict(cfl=float("nan")), "must be finite"),
])
def test_numerics_validation(kwargs, msg):
    with pytest.raises(ValueError, match=msg):
        Numerics(**kwargs)


@pytest.mark.parametrize("kwargs, msg", [
    (dict(sigma=(0.001,)), "exactly two values"),
    (dict(n_protons=0.0), "must be positive"),
    (dict(energy_spread=-1.0), "cannot be negative"),
    (dict(sigma=(0.0, 0.001)), "must be positive"),
    (dict(energy=float("inf")), "must be finite"),
])
def test_beam_validation(kwargs, msg):
    with pytest.raises(ValueError, match=msg):
        Beam(**kwargs)


def test_solver_config_cross_validation():
    with pytest.raises(ValueError, match="outside the x-domain"):
        SolverConfig(geometry=geo(), beam=Beam(energy=100, center=(0.02, 0.0)))
    with pytest.raises(ValueError, match="energy_min must be smaller"):
        SolverConfig(geometry=geo(), beam=Beam(energy=1.0), numerics=Numerics(energy_min=2.0))
    cfg = SolverConfig(geometry=geo(), beam=Beam(energy=100))
    assert cfg.device == DeviceOptions(precision=Precision.FP64, device=0)
    assert cfg.numerics.energy_scheme is EnergyScheme.SSPRK3  # reference default
    assert cfg.numerics.cfl == 0.99


def test_geometry_validation():
    with pytest.raises(ValueError, match="outside the domain"):
        geo(objects=[Box(-0.02, 0.0, -0.01, 0.01, 0.0, 0.01, Material.BONE_COMPACT)])
    with pytest.raises(ValueError, match="may not overlap"):
        geo(objects=[Box(-0.01, 0.0, -0.01, 0.01, 0.0, 0.01, Material.BONE_COMPACT),
                     Box(-0.005, 0.005, -0.01, 0.01, 0.0, 0.01, Material.LUNG)])
    with pytest.raises(ValueError, match="positive integers"):
        Geometry(size=(0.01, 0.01, 0.01), resolution=(10, 10, 2.5))
    with pytest.raises(ValueError, match="minima must be smaller"):
        Box(0.0, 0.0, 0.0, 0.01, 0.0, 0.01, Material.WATER)


def test_box_voxelisation_is_half_open_on_cell_centres():
    g = geo(objects=[Box(0.0, 0.01, -0.01, 0.01, 0.01, 0.02, Material.BONE_COMPACT)])
    grid = Grid.from_geometry(g, np)
    mm = MaterialMap.from_geometry(g, grid, np)
    assert mm.materials == (Material.WATER, Material.BONE_COMPACT)
    inside = ((grid.x >= 0.0) & (grid.x < 0.01))[:, None, None] & \
             ((grid.z >= 0.01) & (grid.z < 0.02))[None, None, :]
    assert np.array_equal(mm.ids == 1, np.broadcast_to(inside, mm.ids.shape))
    rho = density_field(mm, np)
    assert rho.max() == density(Material.BONE_COMPACT) == 1850.0


def test_hu_partition_classes():
    p = HUPartition()
    hu = np.array([[[-1000.0, -200.0, -199.0, 149.9, 150.0, 2000.0]]])
    assert p.classify(hu).tolist() == [[[0, 1, 1, 1, 2, 2]]]
    with pytest.raises(ValueError, match="strictly increasing"):
        HUPartition(cutoffs=(1.0, 1.0))


def test_grid_is_cell_centred():
    grid = Grid.from_geometry(geo(), np)
    assert grid.x[0] == pytest.approx(-0.01 + 0.001)
    assert grid.z[0] == pytest.approx(0.001)
    assert grid.cell_volume == pytest.approx(0.002 ** 3)


def test_stopping_power_host_values_and_range():
    assert stopping_power(Material.WATER, 100.0) == pytest.approx(727.2, rel=0.01)
    with pytest.raises(ValueError, match="outside"):
        stopping_power(Material.WATER, 400.0)


def test_cpu_backend_is_not_this_package():
    from proton_solver.backend import resolve_backend

    with pytest.raises(NotImplementedError):
        resolve_backend(Backend.CPU)


def test_energy_march_argument_errors(monkeypatch):
    from proton_solver.numerics import energy as E

    # moving_active_domain=True goes to the moving-box march, U = None allowed
    cfg = SolverConfig(geometry=geo(), beam=Beam(energy=100),
                       numerics=Numerics(moving_active_domain=True))
    monkeypatch.setattr(E, "_moving_active_energy_march", lambda *a, **k: "moving")
    assert E.energy_march(None, None, None, None, cfg) == "moving"
    energy_march = E.energy_march
    cfg = SolverConfig(geometry=geo(), beam=Beam(energy=100))
    with pytest.raises(ValueError, match="full-grid initial state"):
        energy_march(None, None, None, None, cfg)


# ---------------------------------------------------------------- metrics
def test_range_metrics_on_analytic_curve():
    z = (np.arange(200) + 0.5) * 0.001
    curve = np.where(z < 0.1, 1.0 + 4.0 * z / 0.1, np.maximum(0.0, 5.0 - (z - 0.1) * 1000.0))
    dose = np.broadcast_to(curve, (3, 3, 200)).copy()
    r80, dfw = metrics.range_metrics(dose, z)
    # linear fall-off of 1/mm from 5 at 100 mm: 80% at 101 mm, 20% at 104 mm
    assert r80 == pytest.approx(0.101, abs=2e-4)
    assert dfw == pytest.approx(0.003, abs=2e-4)


def test_central_axis_cell_follows_half_open_rule():
    assert metrics.axis_cell_index(np.array([-1.5, -0.5, 0.5, 1.5])) == 2
    assert metrics.axis_cell_index(np.array([-1.0, 0.0, 1.0])) == 1


def test_gamma_identical_and_shifted_doses():
    rng = np.random.default_rng(0)
    ref = np.zeros((8, 8, 30))
    z = np.arange(30)
    ref[:, :, :] = np.clip(np.minimum(z - 2.0, 27.0 - z) / 5.0, 0.0, 1.0)  # flat top
    assert metrics.gamma_pass_rate(ref, ref, (1e-3, 1e-3, 1e-3)) == 100.0
    # a 2 % dose error everywhere fails 1%/1mm on the plateau (no nearby
    # point has the right dose) but passes 3%/3mm
    assert metrics.gamma_pass_rate(1.02 * ref, ref, (1e-3,) * 3, 1.0, 1.0) < 100.0
    assert metrics.gamma_pass_rate(1.02 * ref, ref, (1e-3,) * 3, 3.0, 3.0) == 100.0
    # a half-voxel shift along z passes 1 mm DTA
    shifted = np.zeros_like(ref)
    shifted[:, :, :] = np.clip(np.minimum(z - 2.5, 26.5 - z) / 5.0, 0.0, 1.0)
    assert metrics.gamma_pass_rate(shifted, ref, (1e-3,) * 3, 1.0, 1.0) == 100.0


def test_gamma_brute_force_point():
    """One evaluated point against an exhaustive gamma by hand."""
    ref = np.zeros((1, 1, 5))
    ref[0, 0, 2] = 1.0
    ref[0, 0, 1] = ref[0, 0, 3] = 0.5
    ev = np.roll(ref, 1, axis=2)  # shifted by one 1 mm voxel
    # at the peak (z=2): evaluated dose at offset +1 mm is 1.0 -> gamma = 1 (distance only)
    g = metrics.gamma_pass_rate(ev, ref, (1e-3,) * 3, 1.0, 1.0, lower_percent_cutoff=90.0)
    assert g == 100.0
    g = metrics.gamma_pass_rate(ev, ref, (1e-3,) * 3, 1.0, 0.8, lower_percent_cutoff=90.0)
    assert g == 0.0


def test_gamma_search_points_outside_the_grid_never_match():
    """pymedphys interpolates the evaluated dose with +inf outside its grid,
    so search points beyond the grid never match.  The host search equals a
    brute-force search with scipy's RegularGridInterpolator filled with +inf
    outside the grid on a case with every voxel on the grid's surface and
    cells (1.5 mm) that the search step (0.1 mm) does not divide."""
    ref = np.ones((3, 3, 3))
    ev = np.full((3, 3, 3), 1.02)
    ev[1] = 1.0
    sp = (1.5e-3, 1.5e-3, 1.5e-3)
    got = metrics.gamma_pass_rate(ev, ref, sp, 1.0, 1.0, 0.0)
    # brute force over the same offsets, skipping search points outside the grid
    setup = metrics._gamma_setup(ev, ref, sp, 1.0, 1.0, 0.0, 10, 1.1)
    _, _, pts, scaled, dist_terms, dd = setup
    from scipy.interpolate import RegularGridInterpolator

    axes = [np.arange(n, dtype=float) for n in ev.shape]
    interp = RegularGridInterpolator(axes, ev, bounds_error=False, fill_value=np.inf)
    n_pass = 0
    for p in pts:
        q = p[None, :] + scaled
        vals = interp(q)
        g2 = ((vals - ref[tuple(p)]) / dd) ** 2 + dist_terms
        n_pass += bool(np.min(g2) <= 1.0)
    assert got == pytest.approx(100.0 * n_pass / len(pts), abs=1e-12)
    # the exact plane passes; the others lie 1.5 mm from it, beyond the reach
    assert got == pytest.approx(100.0 / 3.0)
