"""CT ingestion (geometry.py:43-135 CTScan / HUPartition / Geometry.from_ct_scan,
materials.py:70-97), CPU only.

Pinned by the reference's CT benchmark: the section of its prostate CT that
tests/golden/make_ct_fixture.py cut out (checked there against the source
digest, skin index and outside-body fraction that ct_metrics.json records)
must, refined 6x to 0.5 mm cells and split at 150 HU, give the material
fractions the reference reported (water 0.7472, compact bone 0.2528) on a
60 x 60 x 300 grid of 30 x 30 x 150 mm.  The orientation options, key
names and validation messages follow the reference.  The march on this
geometry is compared with the reference's dose on the GPU
(tests/test_gpu_ct.py).
"""

import numpy as np
import pytest

from helpers import ct_config, ct_golden, ct_section_npz
from proton_solver import CTScan, Geometry, HUPartition, Material
from proton_solver.numerics.grid import Grid
from proton_solver.physics.materials import MaterialMap


def test_reference_section_material_fractions(tmp_path):
    g = ct_golden()
    cfg = ct_config(ct_section_npz(tmp_path))
    geo = cfg.geometry
    assert geo.resolution == (60, 60, 300)
    assert geo.size == pytest.approx((0.03, 0.03, 0.15), rel=1e-12)
    grid = Grid.from_geometry(geo, np)
    mm = MaterialMap.from_geometry(geo, grid, np)
    assert mm.materials == (Material.WATER, Material.BONE_COMPACT)
    frac = {m.name: float((mm.ids == i).mean()) for i, m in enumerate(mm.materials)}
    for name, value in g["material_fractions"].items():
        assert frac[name] == pytest.approx(value, abs=1e-12)
    assert float(1.0 - geo.ct_scan.body_mask.mean()) == pytest.approx(
        g["section"]["outside_body_voxels_fraction"], abs=1e-12)


def test_partition_bins_are_half_open():
    p = HUPartition()          # (-200, 150): lung | water | bone
    hu = np.array([-1000, -200.0001, -200, 0, 149.999, 150, 3000], float).reshape(1, 1, 7)
    assert p.classify(hu).ravel().tolist() == [0, 0, 1, 1, 1, 2, 2]
    with pytest.raises(ValueError, match="one more material"):
        HUPartition(cutoffs=(0.0,), materials=(Material.WATER,))
    with pytest.raises(ValueError, match="strictly increasing"):
        HUPartition(cutoffs=(1.0, 1.0), materials=(Material.WATER,) * 3)
    with pytest.raises(ValueError, match="non-finite"):
        p.classify(np.full((1, 1, 2), np.nan))


def test_orientation_keys_and_mask(tmp_path):
    rng = np.random.default_rng(0)
    hu = rng.integers(-1000, 1500, size=(3, 4, 5)).astype(np.int16)
    mask = rng.random((3, 4, 5)) > 0.3
    f = tmp_path / "a.npz"
    np.savez(f, hu=hu, body_mask=mask)
    ct = CTScan(f, (1e-3, 2e-3, 3e-3), axis_order=(2, 0, 1), flip_axes=(True, False, False))
    want = np.flip(np.transpose(hu, (2, 0, 1)), 0)
    assert np.array_equal(ct.hu, want) and ct.shape == (5, 3, 4)
    assert np.array_equal(ct.body_mask, np.flip(np.transpose(mask, (2, 0, 1)), 0))
    assert ct.spacing == (3e-3, 1e-3, 2e-3)
    g = tmp_path / "b.npz"
    np.savez(g, hu)                      # a single arr_0
    assert np.array_equal(CTScan(g, (1e-3,) * 3).hu, hu)
    # outside the body the background material replaces the HU class
    geo = Geometry.from_ct_scan(f, source_spacing_m=(1e-3, 1e-3, 1e-3),
                                background_material=Material.AIR)
    grid = Grid.from_geometry(geo, np)
    mm = MaterialMap.from_geometry(geo, grid, np)
    air = mm.materials.index(Material.AIR)
    assert np.all(mm.ids[~mask] == air)
    cls = HUPartition().classify(hu)
    inside = [mm.materials.index(HUPartition().materials[c]) for c in cls[mask]]
    assert np.array_equal(mm.ids[mask], inside)


def test_ct_validation(tmp_path):
    f = tmp_path / "c.npz"
    np.savez(f, other=np.zeros((2, 2, 2)), more=np.zeros(2))
    with pytest.raises(ValueError, match="'HU', 'hu', or single 'arr_0'"):
        CTScan(f, (1e-3,) * 3)
    np.savez(f, HU=np.zeros((2, 2, 2)), body_mask=np.ones((2, 2, 3), bool))
    with pytest.raises(ValueError, match="body_mask shape"):
        CTScan(f, (1e-3,) * 3)
    np.savez(f, HU=np.zeros((2, 2, 2)))
    with pytest.raises(ValueError, match="permutation"):
        CTScan(f, (1e-3,) * 3, axis_order=(0, 0, 1))
    with pytest.raises(ValueError, match="spacing"):
        CTScan(f, (1e-3, 0.0, 1e-3))
    ct = CTScan(f, (1e-3,) * 3)
    with pytest.raises(ValueError, match="resolution must match"):
        Geometry(size=(0.002, 0.002, 0.003), resolution=(2, 2, 3), ct_scan=ct)
