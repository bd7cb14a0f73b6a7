"""Batched-beamlet march on an MI355X, through the C ABI (m1_batch_*).

A small phantom (water with a bone and a lung ellipsoid, 48 x 48 x 72 cells
at 1 mm) and three beamlets on 32 x 32 boxes.  Checked:
  * each beamlet of the batch equals the same box marched alone by the
    single-box FP32-bulk march (same kernels, per-box tile sets and CFL);
  * the phantom dose is the sum of the beamlet doses;
  * against the plain FP64 march of each beamlet on the whole phantom (which
    is bitwise the CPU oracle): max-dose error <= 0.5 % and
    gamma(1 %/1 mm) >= 99.5 %, the acceptance of the mixed-precision march;
  * the argument checks of the batch entry points.
"""

import numpy as np
import pytest

from proton_solver import (
    DeviceOptions, EnergyScheme, Material, Numerics, Physics, Precision, Scattering,
)
from proton_solver import _native, metrics
from proton_solver.batch import (
    LUNG_026, Beamlet, Ellipsoid, batch_march, beamlet_boxes, box_problem, ellipsoid_phantom,
    full_problem, prepare_batch,
)
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


class _Cfg:
    """The parts of a SolverConfig the batch reads."""

    def __init__(self, precision=Precision.FP32_BULK, tile_threshold=1e-10,
                 scheme=EnergyScheme.IF_SSPRK3, physics=None):
        self.physics = physics or Physics(scattering=Scattering.WENTZEL, nuclear=True)
        self.numerics = Numerics(energy_scheme=scheme, cfl=0.9, max_energy_step=0.5,
                                 energy_min=1.0)
        self.device = DeviceOptions(precision=precision, tile_threshold=tile_threshold)


def small_workload():
    ell = [Ellipsoid((0.006, 0.0, 0.030), (0.006, 0.008, 0.010), Material.BONE_COMPACT),
           Ellipsoid((-0.008, 0.004, 0.018), (0.007, 0.007, 0.012), LUNG_026)]
    ph = ellipsoid_phantom((48, 48, 72), (0.001,) * 3, ell)
    beamlets = [Beamlet(energy=45.0, sigma=(0.002, 0.002), center=(-0.0043, 0.0021)),
                Beamlet(energy=62.0, sigma=(0.003, 0.003), center=(0.0031, -0.0028)),
                Beamlet(energy=55.0, sigma=(0.0025, 0.0025), center=(0.0004, 0.0057))]
    boxes = beamlet_boxes(ph, beamlets, lateral=(32, 32))
    return ph, beamlets, boxes


@pytest.fixture(scope="module")
def batch_result():
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg()
    say("batch boxes", boxes)
    res = batch_march(ph, beamlets, cfg, boxes=boxes)
    say("batch steps", res.steps.tolist())
    return ph, beamlets, boxes, cfg, res


def test_batch_beamlets_equal_single_box_marches(batch_result):
    ph, beamlets, boxes, cfg, res = batch_result
    num = _native.numerics_struct(cfg)
    ctx = _native.Context(0)
    for b, (box, bl) in enumerate(zip(boxes, beamlets)):
        grid, mm, inflow, beam = box_problem(ph, box, bl)
        ctx.set_problem(grid, mm, inflow, beam)
        E, _ = ctx.march(num)
        single = ctx.get_dose()
        got = res.beamlet_dose(b)
        diff = np.abs(got - single).max() / single.max()
        say("beamlet", b, "steps", res.steps[b], len(E) - 1, "max|d|/max", diff,
            "bitwise", bool(np.array_equal(got, single)))
        assert res.steps[b] == len(E) - 1
        assert diff <= 1e-6


def test_batch_phantom_dose_is_the_sum_of_beamlet_doses(batch_result):
    ph, beamlets, boxes, cfg, res = batch_result
    total = np.zeros(ph.shape)
    for b, (x0, y0, nx, ny, nz) in enumerate(boxes):
        total[x0:x0 + nx, y0:y0 + ny, :nz] += res.beamlet_dose(b)
    assert np.array_equal(total, res.dose)


def test_batch_against_fp64_full_phantom_march(batch_result):
    ph, beamlets, boxes, cfg, res = batch_result
    num64 = _native.numerics_struct(cfg, _native.FP64)
    ctx = _native.Context(0)
    spacing = ph.spacing
    for b, (box, bl) in enumerate(zip(boxes, beamlets)):
        grid, mm, inflow, beam = full_problem(ph, bl)
        ctx.set_problem(grid, mm, inflow, beam)
        ctx.set_state(None)
        E, _ = ctx.march(num64)
        ref = ctx.get_dose()
        got = np.zeros(ph.shape)
        x0, y0, nx, ny, nz = box
        got[x0:x0 + nx, y0:y0 + ny, :nz] = res.beamlet_dose(b)
        md = metrics.maximum_dose_relative_difference(got, ref)
        gam = metrics.gamma_pass_rate(got, ref, spacing, 1.0, 1.0, 10.0)
        outside = ref.copy()
        outside[x0:x0 + nx, y0:y0 + ny, :nz] = 0.0
        say("beamlet", b, "E", bl.energy, "steps", res.steps[b], "vs fp64", len(E) - 1,
            "max-dose err", md, "gamma", gam, "fp64 dose outside box / max",
            outside.max() / ref.max())
        assert abs(md) <= 5e-3
        assert gam >= 99.5


def test_batch_tile_skipping_against_every_tile():
    ph, beamlets, boxes = small_workload()
    h_skip, _ = prepare_batch(ph, beamlets, _Cfg(), boxes)
    h_skip.reset_stats()
    d1, s1 = h_skip.march(_native.numerics_struct(_Cfg()))
    ev, full = h_skip.tile_stats()
    d0, s0 = h_skip.march(_native.numerics_struct(_Cfg(tile_threshold=0.0)))
    err = metrics.max_abs_difference_over_max(d1, d0)
    say("batch tiles evaluated", ev, "of", full, "max|dD|/maxD", err, s1.tolist(), s0.tolist())
    assert 0 < ev < full
    assert err < 5e-4


def test_batch_argument_checks():
    ph, beamlets, boxes = small_workload()
    with pytest.raises(NotImplementedError, match="FP32-bulk"):
        batch_march(ph, beamlets, _Cfg(precision=Precision.FP64), boxes=boxes)
    h, _ = prepare_batch(ph, beamlets, _Cfg(), boxes)
    with pytest.raises(NotImplementedError, match="FP32-bulk"):
        h.march(_native.numerics_struct(_Cfg(), _native.FP64))
    with pytest.raises(NotImplementedError, match="IF-SSPRK3"):
        h.march(_native.numerics_struct(_Cfg(scheme=EnergyScheme.SSPRK3)))
    with pytest.raises(ValueError, match="no march"):
        h.beamlet_dose(0)
    with pytest.raises(ValueError, match="out of range"):
        h.march(_native.numerics_struct(_Cfg()))
        h.beamlet_dose(7)
    bad = list(boxes)
    bad[0] = (40, 0, 32, 32, 64)
    with pytest.raises(ValueError, match="outside the phantom"):
        prepare_batch(ph, beamlets, _Cfg(), bad)
    # the library's own check, past the host-side one
    profiles = [np.ones((32, 32))] * 3
    with pytest.raises(ValueError, match="outside the phantom"):
        h.set_problem(ph.ids, ph.spacing, ph.materials, bad, profiles,
                      [bl.beam() for bl in beamlets])
