"""Geometry validation and CT ingestion (geometry.py:10-271), CPU only.

Every check of the reference's Box, HUPartition, CTScan and Geometry raises
ValueError with the reference's message; the overlap sweep agrees with the
all-pairs rule on random boxes; CT orientation is a permutation of axes
followed by flips, checked against explicit index arithmetic.
"""

import math

import numpy as np
import pytest

from proton_solver import Box, CTScan, Geometry, HUPartition, Material

W = Material.WATER


def test_box_checks():
    with pytest.raises(ValueError, match="Box coordinates must be finite."):
        Box(0, math.nan, 0, 1, 0, 1, W)
    with pytest.raises(ValueError, match="Box minima must be smaller than maxima."):
        Box(0, 1, 0, 1, 1, 1, W)
    b = Box(-1, 1, -2, 2, 0, 3, W)
    assert b.intervals() == ((-1, 1), (-2, 2), (0, 3))


def test_hu_partition_checks_and_bins():
    with pytest.raises(ValueError, match="one more material than cutoffs"):
        HUPartition(cutoffs=(0.0,), materials=(W,))
    with pytest.raises(ValueError, match="must be Material values"):
        HUPartition(cutoffs=(0.0,), materials=(W, "bone"))
    with pytest.raises(ValueError, match="HU cutoffs must be finite."):
        HUPartition(cutoffs=(0.0, math.inf), materials=(W, W, W))
    with pytest.raises(ValueError, match="strictly increasing"):
        HUPartition(cutoffs=(5.0, 5.0), materials=(W, W, W))
    p = HUPartition()
    hu = np.array([-1000, -200.5, -200, 149.9, 150, 3000], dtype=float).reshape(1, 2, 3)
    assert p.classify(hu).ravel().tolist() == [0, 0, 1, 1, 2, 2]
    with pytest.raises(ValueError, match="numeric 3D array"):
        p.classify(np.zeros((2, 2)))
    with pytest.raises(ValueError, match="non-finite"):
        p.classify(np.full((1, 1, 1), np.nan))


def _brute_overlap(boxes):
    return any(a.overlaps(b) for i, a in enumerate(boxes) for b in boxes[i + 1:])


def test_overlap_sweep_agrees_with_all_pairs():
    rng = np.random.default_rng(7)
    size = (0.1, 0.1, 0.1)
    n_overlapping = 0
    for trial in range(300):
        boxes = []
        for _ in range(int(rng.integers(1, 7))):
            lo = rng.uniform([-0.05, -0.05, 0.0], [0.03, 0.03, 0.08])
            hi = lo + rng.uniform(0.01, 0.05, 3)
            hi = np.minimum(hi, [0.05, 0.05, 0.1])
            # snap to a coarse grid so that touching faces occur
            lo, hi = np.round(lo, 2), np.round(hi, 2)
            if np.any(hi <= lo):
                continue
            boxes.append(Box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], W))
        expect = _brute_overlap(boxes)
        n_overlapping += expect
        if expect:
            with pytest.raises(ValueError, match="Geometry objects may not overlap."):
                Geometry(size=size, resolution=(10, 10, 10), objects=boxes)
        else:
            Geometry(size=size, resolution=(10, 10, 10), objects=boxes)
    assert 20 < n_overlapping < 280
    # half-open: touching faces do not overlap
    Geometry(size=size, resolution=(10, 10, 10),
             objects=[Box(-0.05, 0.0, -0.05, 0.05, 0, 0.1, W),
                      Box(0.0, 0.05, -0.05, 0.05, 0, 0.1, W)])


def test_geometry_checks():
    with pytest.raises(ValueError, match="exactly three values"):
        Geometry(size=(0.1, 0.1), resolution=(1, 1, 1))
    with pytest.raises(ValueError, match="finite and positive"):
        Geometry(size=(0.1, -0.1, 0.1), resolution=(1, 1, 1))
    with pytest.raises(ValueError, match="positive integers"):
        Geometry(size=(0.1, 0.1, 0.1), resolution=(1, 1.0, 1))
    with pytest.raises(ValueError, match="extends outside the domain"):
        Geometry(size=(0.1, 0.1, 0.1), resolution=(1, 1, 1),
                 objects=[Box(-0.06, 0.0, -0.05, 0.05, 0.0, 0.1, W)])
    with pytest.raises(ValueError, match="extends outside the domain"):
        Geometry(size=(0.1, 0.1, 0.1), resolution=(1, 1, 1),
                 objects=[Box(-0.05, 0.0, -0.05, 0.05, -0.01, 0.1, W)])
    g = Geometry(size=(0.02, 0.04, 0.05), resolution=(4, 8, 25))
    assert g.spacing == pytest.approx((0.005, 0.005, 0.002)) and g.shape == (4, 8, 25)


def _write_ct(path, hu, mask=None, key="HU"):
    arrays = {key: hu}
    if mask is not None:
        arrays["body_mask"] = mask
    np.savez(path, **arrays)
    return path


def test_ct_orientation_is_permutation_then_flips(tmp_path):
    rng = np.random.default_rng(3)
    hu = rng.integers(-1000, 2000, size=(3, 4, 5)).astype(np.int16)
    mask = hu > 0
    path = _write_ct(tmp_path / "ct.npz", hu, mask)
    order, flips = (2, 0, 1), (True, False, True)
    scan = CTScan(path, (0.001, 0.002, 0.003), order, flips)
    assert scan.shape == (5, 3, 4) and scan.spacing == (0.003, 0.001, 0.002)
    for i in range(5):
        for j in range(3):
            for k in range(4):
                src = [0, 0, 0]
                # solver axis a holds source axis order[a], reversed when flipped
                src[order[0]] = 4 - i
                src[order[1]] = j
                src[order[2]] = 3 - k
                assert scan.hu[i, j, k] == hu[tuple(src)]
                assert scan.body_mask[i, j, k] == mask[tuple(src)]
    assert scan.hu.flags.c_contiguous and scan.body_mask.dtype == bool


def test_ct_checks(tmp_path):
    hu = np.zeros((2, 3, 4))
    good = _write_ct(tmp_path / "a.npz", hu)
    with pytest.raises(ValueError, match="three positive finite metres"):
        CTScan(good, (0.001, 0.0, 0.001))
    with pytest.raises(ValueError, match="permutation of"):
        CTScan(good, (0.001,) * 3, axis_order=(0, 0, 1))
    with pytest.raises(ValueError, match="three booleans"):
        CTScan(good, (0.001,) * 3, flip_axes=(0, 1, 0))
    np.savez(tmp_path / "b.npz", volume=hu)
    with pytest.raises(ValueError, match="single 'arr_0' array"):
        CTScan(tmp_path / "b.npz", (0.001,) * 3)
    np.savez(tmp_path / "c.npz", hu)
    assert CTScan(tmp_path / "c.npz", (0.001,) * 3).shape == (2, 3, 4)
    assert CTScan(_write_ct(tmp_path / "d.npz", hu, key="hu"), (0.001,) * 3).body_mask is None
    with pytest.raises(ValueError, match="numeric 3D array"):
        CTScan(_write_ct(tmp_path / "e.npz", np.zeros((2, 3))), (0.001,) * 3)
    with pytest.raises(ValueError, match="non-finite"):
        CTScan(_write_ct(tmp_path / "f.npz", np.full((1, 1, 1), np.inf)), (0.001,) * 3)
    with pytest.raises(ValueError, match="body_mask shape"):
        CTScan(_write_ct(tmp_path / "g.npz", hu, np.ones((2, 3, 3), bool)), (0.001,) * 3)


def test_geometry_from_ct_scan_and_its_checks(tmp_path):
    path = _write_ct(tmp_path / "ct.npz", np.zeros((2, 3, 4)))
    g = Geometry.from_ct_scan(path, source_spacing_m=(0.001, 0.002, 0.003))
    assert g.resolution == (2, 3, 4) and g.size == pytest.approx((0.002, 0.006, 0.012))
    assert isinstance(g.hu_partition, HUPartition)
    scan = g.ct_scan
    with pytest.raises(ValueError, match="cannot be combined"):
        Geometry(size=g.size, resolution=g.resolution, ct_scan=scan,
                 objects=[Box(-0.001, 0.001, -0.003, 0.003, 0, 0.012, W)])
    with pytest.raises(ValueError, match="match the CT array shape"):
        Geometry(size=g.size, resolution=(2, 3, 5), ct_scan=scan)
    with pytest.raises(ValueError, match="match the CT spacing"):
        Geometry(size=(0.004, 0.006, 0.012), resolution=(2, 3, 4), ct_scan=scan)
    with pytest.raises(ValueError, match="does not support CT"):
        Geometry(size=g.size, resolution=g.resolution, ct_scan=scan, view_geometry=True)
