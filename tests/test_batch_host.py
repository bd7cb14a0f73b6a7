"""Host side of the batched-beamlet march (CPU only, no device calls).

Phantom voxelisation, the seeded workload, beamlet boxes and their depth
rule, per-box inflow profiles, and the density-scaled material (lung at
0.26 g/cm^3) that the workload uses.
"""

import math

import numpy as np
import pytest
from scipy.integrate import quad

from proton_solver import Material
from proton_solver.batch import (
    FIVE_MATERIALS, LUNG_026, Beamlet, Ellipsoid, air_free_beamlets, batch_workload,
    beamlet_boxes, box_depths, box_problem, combined_workload, ellipsoid_phantom, full_problem,
)
from proton_solver.physics.materials import ScaledMaterial, density, radiation_length
from proton_solver.physics.nuclear import load_nuclear_table
from proton_solver.physics.scattering import load_transport_table
from proton_solver.physics.stopping_power import load_stopping_power_table, stopping_power


# ------------------------------------------------------------ materials
def test_scaled_material_tables_and_constants():
    base = Material.LUNG
    L = LUNG_026
    assert density(L) == pytest.approx(260.0, rel=1e-15)
    assert radiation_length(L) == pytest.approx(4.0 * radiation_length(base), rel=1e-15)
    for load in (load_stopping_power_table, load_transport_table, load_nuclear_table):
        E, v = load(L)
        E0, v0 = load(base)
        assert np.array_equal(E, E0)
        assert np.array_equal(v, 0.25 * v0)  # power-of-two scale: exact
    for e in (3.0, 50.0, 181.7):
        assert stopping_power(L, e) == pytest.approx(0.25 * stopping_power(base, e), rel=1e-14)


def test_scaled_material_leaves_optical_depths_unchanged():
    """tau = int Sigma/S dE is independent of density at fixed composition."""
    from oracle import oracle as O

    m0, k0 = O.material_struct(Material.LUNG)
    m1, k1 = O.material_struct(LUNG_026)
    for e, d in ((120.0, 0.5), (30.0, 0.25)):
        for model in ("WENTZEL", "ROSSI"):
            t0 = O.scattering_depth(m0, model, e, d)
            t1 = O.scattering_depth(m1, model, e, d)
            assert t1 == pytest.approx(t0, rel=1e-13)
        assert O.nuclear_depth(m1, e, d) == pytest.approx(O.nuclear_depth(m0, e, d), rel=1e-13)


def test_scaled_material_validation():
    with pytest.raises(ValueError, match="density_scale"):
        ScaledMaterial(Material.LUNG, 0.0)
    with pytest.raises(ValueError, match="base"):
        ScaledMaterial("lung", 0.5)


# ------------------------------------------------------------ phantoms
def test_ellipsoid_voxelisation_by_cell_centre():
    ell = [Ellipsoid((0.002, -0.003, 0.010), (0.004, 0.006, 0.005), Material.BONE_COMPACT),
           Ellipsoid((0.0, 0.0, 0.012), (0.003, 0.003, 0.003), LUNG_026)]
    ph = ellipsoid_phantom((20, 24, 30), (0.001, 0.001, 0.001), ell)
    assert ph.materials == (Material.WATER, Material.BONE_COMPACT, LUNG_026)
    g = ph.grid()
    rng = np.random.default_rng(0)
    for _ in range(400):
        i, j, k = (int(rng.integers(n)) for n in ph.shape)
        want = 0
        for idx, e in enumerate(ell, start=1):
            r = sum(((c - q) / a) ** 2 for c, q, a in zip((g.x[i], g.y[j], g.z[k]), e.center,
                                                            e.semi_axes))
            if r <= 1.0:
                want = idx  # later ellipsoids overwrite earlier ones
        assert ph.ids[i, j, k] == want


def test_batch_workload_is_seeded_and_in_range():
    ph, bl = batch_workload()
    ph2, bl2 = batch_workload()
    assert np.array_equal(ph.ids, ph2.ids) and bl == bl2
    assert ph.shape == (200, 200, 300)
    assert set(ph.materials) == {Material.WATER, Material.BONE_COMPACT, LUNG_026}
    assert len(bl) == 256
    E = np.array([b.energy for b in bl])
    s = np.array([b.sigma[0] for b in bl])
    c = np.array([b.center for b in bl])
    assert 70.0 <= E.min() and E.max() <= 230.0
    assert 0.002 <= s.min() and s.max() <= 0.006
    assert np.abs(c).max() <= 0.030
    # uniform draws: the sample spans most of each range
    assert E.max() - E.min() > 140.0 and s.max() - s.min() > 0.0035


def test_combined_workload_is_seeded_and_in_range():
    """256 x 256 x 200 phantom, 40 ellipsoids over the five tabulated
    materials, 1024 beamlets with the 256-beamlet distributions."""
    ph, bl = combined_workload()
    ph2, bl2 = combined_workload()
    assert np.array_equal(ph.ids, ph2.ids) and bl == bl2
    assert ph.shape == (256, 256, 200) and ph.spacing == (0.001, 0.001, 0.001)
    assert set(ph.materials) == set(FIVE_MATERIALS) and ph.materials[0] is Material.WATER
    counts = np.bincount(ph.ids.ravel(), minlength=len(ph.materials))
    assert counts.min() > 0 and counts[0] > 0.9 * ph.ids.size
    assert len(bl) == 1024
    E = np.array([b.energy for b in bl])
    c = np.array([b.center for b in bl])
    assert 70.0 <= E.min() and E.max() <= 230.0 and np.abs(c).max() <= 0.030
    # the first 256 beamlets differ from the 256-beamlet workload's (other seed)
    assert [b.energy for b in bl[:256]] != [b.energy for b in batch_workload()[1]]


def test_air_free_beamlets_of_the_combined_workload():
    """131 of the 1024 beamlets have no air within 48 cells of their entry
    column over their box depth; a phantom without air keeps every beamlet."""
    ph, bl = combined_workload()
    boxes = beamlet_boxes(ph, bl)
    idx = air_free_beamlets(ph, bl, boxes)
    assert len(idx) == 131 and idx[:6] == [1, 3, 19, 21, 23, 29]
    air = ph.ids == ph.materials.index(Material.AIR)
    for b in idx[:20]:
        ix = int(np.floor(bl[b].center[0] / 0.001)) + 128
        iy = int(np.floor(bl[b].center[1] / 0.001)) + 128
        assert not air[max(ix - 48, 0):ix + 48, max(iy - 48, 0):iy + 48, :boxes[b][4]].any()
    # the rejected ones do have air in that footprint (the first five checked)
    rejected = [b for b in range(20) if b not in idx][:5]
    for b in rejected:
        ix = int(np.floor(bl[b].center[0] / 0.001)) + 128
        iy = int(np.floor(bl[b].center[1] / 0.001)) + 128
        assert air[max(ix - 48, 0):ix + 48, max(iy - 48, 0):iy + 48, :boxes[b][4]].any()
    ph4, bl4 = batch_workload(n_beamlets=8)
    assert air_free_beamlets(ph4, bl4, beamlet_boxes(ph4, bl4)) == list(range(8))


def test_cfl_step_in_air_is_about_a_thousandth_of_water():
    """Energy step of the CFL bound at the isotropic speed 1/sqrt(3) on
    1 mm cells, dE = cfl / sum_axes(a / (S dx)), per material: air's linear
    stopping power is 1/934 of water's at 70-230 MeV, so one active
    (or, in the whole-phantom bound, present) air cell shrinks the step by
    that factor."""
    iso = 1.0 / math.sqrt(3.0)
    for e in (70.0, 150.0, 230.0):
        step = {m: 0.9 / (3 * iso / (stopping_power(m, e) * 0.001)) for m in FIVE_MATERIALS}
        ratio = step[Material.WATER] / step[Material.AIR]
        assert 930.0 < ratio < 940.0, (e, ratio)
        assert step[Material.BONE_COMPACT] > step[Material.WATER] > step[Material.AIR]


# ------------------------------------------------------------ boxes
def _csda_range(material, e0, e_min):
    val, _ = quad(lambda e: 1.0 / stopping_power(material, e), e_min, e0, limit=400)
    return val


def test_box_depth_in_water_is_the_csda_range():
    ph = ellipsoid_phantom((16, 16, 200), (0.001,) * 3, [])
    for e0 in (60.0, 105.0, 140.0):
        R = _csda_range(Material.WATER, e0, 1.0)
        d = box_depths(ph, [(0, 0, 16, 16)], [e0], 1.0, guard_cells=0)[0]
        # first whole cell at which the range is used up (0.01 mm table slack)
        assert d == math.ceil(R / 0.001 - 1e-2) or d == math.ceil(R / 0.001)


def test_box_depth_follows_the_least_stopping_material():
    """A lung slab anywhere in the box's cross-section stretches the range by
    the distance the slab saves; bone elsewhere in the slice does not shorten it."""
    shape = (16, 16, 250)
    lung = Ellipsoid((0.0, 0.0, 0.030), (0.004, 0.004, 0.010), LUNG_026)  # z 20..40 mm
    bone = Ellipsoid((0.005, 0.005, 0.060), (0.002, 0.002, 0.010), Material.BONE_COMPACT)
    ph_w = ellipsoid_phantom(shape, (0.001,) * 3, [])
    ph_l = ellipsoid_phantom(shape, (0.001,) * 3, [lung, bone])
    d_w = box_depths(ph_w, [(0, 0, 16, 16)], [150.0], 1.0, guard_cells=0)[0]
    d_l = box_depths(ph_l, [(0, 0, 16, 16)], [150.0], 1.0, guard_cells=0)[0]
    # the lung column is 20 mm of 0.26 g/cm^3: it saves ~0.74 * 20 = 14.8 mm of water
    assert 13 <= d_l - d_w <= 17


def test_beamlet_boxes_contain_the_entry_point_and_fit_the_phantom():
    ph, bl = batch_workload(n_beamlets=24)
    boxes = beamlet_boxes(ph, bl)
    g = ph.grid()
    for b, (x0, y0, nx, ny, nz) in zip(bl, boxes):
        assert (nx, ny) == (64, 64)
        assert 0 <= x0 and x0 + nx <= 200 and 0 <= y0 and y0 + ny <= 200 and 4 <= nz <= 300
        assert nz % 32 == 0 or nz == 300
        # the entry point lies in the cell just above the box centre plane
        cx = g.x[x0] - 0.0005 + 0.5 * nx * 0.001
        cy = g.y[y0] - 0.0005 + 0.5 * ny * 0.001
        ox, oy = b.center[0] - cx, b.center[1] - cy
        assert -1e-12 <= ox < 0.001 + 1e-12 and -1e-12 <= oy < 0.001 + 1e-12
        # 5 sigma cut-off disc inside the box
        assert 0.5 * nx * 0.001 - max(ox, oy) >= b.cutoff_sigma * b.sigma[0]


def test_box_inflow_is_the_full_phantom_inflow_restricted_to_the_box():
    ph = ellipsoid_phantom((60, 60, 40), (0.001,) * 3, [])
    bl = Beamlet(energy=80.0, sigma=(0.003, 0.003), center=(0.0042, -0.0071))
    box = beamlet_boxes(ph, [bl], lateral=(36, 36))[0]
    x0, y0, nx, ny, nz = box
    _, _, inflow_b, beam_b = box_problem(ph, box, bl)
    _, _, inflow_f, _ = full_problem(ph, bl)
    sub = np.asarray(inflow_f.spatial)[x0:x0 + nx, y0:y0 + ny]
    assert np.asarray(inflow_f.spatial).sum() == pytest.approx(sub.sum(), rel=1e-14)
    np.testing.assert_allclose(np.asarray(inflow_b.spatial), sub, rtol=1e-9, atol=0)


def test_batch_tile_shapes_listed_by_the_library():
    """The library reports the tile shapes its batched kernels were built for
    (no GPU call); the default is 16^3 and every shape maps 256 threads."""
    from proton_solver import _native

    shapes = _native.batch_tile_shapes()
    assert shapes[0] == (16, 16, 16) and len(shapes) >= 2
    assert len(set(shapes)) == len(shapes)
    for tx, ty, tz in shapes:
        assert ty * tz == 256 and tx * ty <= 256 and min(tx, ty, tz) >= 8


def test_autotune_sample_spans_the_energies():
    from proton_solver.autotune import sample_beamlets
    from proton_solver.batch import random_beamlets

    bl = random_beamlets(np.random.default_rng(3), 40)
    sub = sample_beamlets(bl, 6)
    energies = sorted(b.energy for b in bl)
    got = sorted(b.energy for b in sub)
    assert len(sub) == 6 and got[0] == energies[0] and got[-1] == energies[-1]
    assert sample_beamlets(bl, 100) and len(sample_beamlets(bl, 100)) == 40
