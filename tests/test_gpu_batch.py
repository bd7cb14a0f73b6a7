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
    DeviceOptions, EnergyScheme, Material, Numerics, Physics, Precision, Scattering, WaveSpeeds,
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
                 scheme=EnergyScheme.IF_SSPRK3, physics=None, vacuum_aware_cfl=False,
                 batch_cfl="phantom"):
        self.physics = physics or Physics(scattering=Scattering.WENTZEL, nuclear=True)
        self.numerics = Numerics(energy_scheme=scheme, cfl=0.9, max_energy_step=0.5,
                                 energy_min=1.0, vacuum_aware_cfl=vacuum_aware_cfl)
        self.device = DeviceOptions(precision=precision, tile_threshold=tile_threshold,
                                    batch_cfl=batch_cfl)


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
    """The batch with each beamlet's CFL step bounded by its box alone (the
    single-box march's rule)."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg(batch_cfl="box")
    say("batch boxes", boxes)
    res = batch_march(ph, beamlets, cfg, boxes=boxes)
    say("batch steps", res.steps.tolist())
    return ph, beamlets, boxes, cfg, res


@pytest.fixture(scope="module")
def batch_phantom_cfl():
    """The batch with the whole-phantom CFL bound (the default)."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg()
    res = batch_march(ph, beamlets, cfg, boxes=boxes)
    say("batch steps, whole-phantom CFL", res.steps.tolist())
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


def test_batch_against_fp64_full_phantom_march(batch_phantom_cfl):
    """With the whole-phantom CFL bound the batch's energy grid is at least as
    fine as that of the FP64 march on the whole phantom: outside the box the
    batch counts the zero state (isotropic speeds), an upper bound of the
    near-vacuum cells there."""
    ph, beamlets, boxes, cfg, res = batch_phantom_cfl
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
        assert int(res.steps[b]) >= len(E) - 1 - 1


def test_batch_tile_skipping_against_every_tile():
    """Plain CFL (every cell, as the reference's energy_march): zeroing the
    sub-threshold tails changes which near-vacuum cell sets the CFL maximum,
    so the energy grids differ slightly (one step here)."""
    ph, beamlets, boxes = small_workload()
    h_skip, _ = prepare_batch(ph, beamlets, _Cfg(), boxes)
    h_skip.reset_stats()
    d1, s1 = h_skip.march(_native.numerics_struct(_Cfg()))
    ev, full = h_skip.tile_stats()
    d0, s0 = h_skip.march(_native.numerics_struct(_Cfg(tile_threshold=0.0)))
    err = metrics.max_abs_difference_over_max(d1, d0)
    md = metrics.maximum_dose_relative_difference(d1, d0)
    say("batch tiles evaluated", ev, "of", full, "max|dD|/maxD", err, "max-dose", md,
        s1.tolist(), s0.tolist())
    assert 0 < ev < full
    assert err < 2e-3 and abs(md) < 1.5e-3


def test_batch_tile_skipping_is_cfl_neutral_with_the_vacuum_aware_cfl():
    """Vacuum-aware CFL (energy.py:459-537): only cells within 3 cells of an
    active cell set the step, so skipping tiles no longer moves the step
    count; what remains is the speed of active cells near the threshold,
    whose state feels the dropped sub-threshold tails (1.9e-4 of the maximum
    observed here against the march over every tile, recording mode, same
    activity threshold)."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg(vacuum_aware_cfl=True)
    num = _native.numerics_struct(cfg)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes)
    h.reset_stats()
    d1, s1 = h.march(num)
    ev, full = h.tile_stats()
    h.record(True, full_domain=True)
    h.reset_stats()
    d0, s0 = h.march(num)
    ev0, _ = h.tile_stats()
    h.record(False)
    err = metrics.max_abs_difference_over_max(d1, d0)
    say("vacuum-aware CFL: tiles", ev, "vs every tile", ev0, "of", full, "max|dD|/maxD", err,
        "steps", s1.tolist(), s0.tolist())
    assert ev0 == full and ev < full
    assert np.array_equal(s1, s0)
    assert err < 5e-4
    # fewer steps than the plain CFL, whose maximum near-vacuum cells set
    _, s_plain = h.march(_native.numerics_struct(_Cfg()))
    say("steps with the plain CFL", s_plain.tolist())
    assert (s1 <= s_plain).all()


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


def test_persistent_launch_is_bitwise_the_per_tile_launch():
    """Stage kernels on a persistent grid with per-XCD tile queues compute the
    same tiles with the same arithmetic as one workgroup per tile."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg()
    num = _native.numerics_struct(cfg)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes)
    d0, s0 = h.march(num)
    h.set_persistent(True)
    d1, s1 = h.march(num)
    h.set_persistent(False)
    say("persistent launch: steps", s1.tolist(), "bitwise", bool(np.array_equal(d0, d1)))
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1)


@pytest.mark.parametrize("threshold,batch_cfl,wave", [
    (1e-10, "phantom", WaveSpeeds.EXACT), (1e-6, "phantom", WaveSpeeds.EXACT),
    (1e-6, "box", WaveSpeeds.EXACT), (1e-6, "phantom", WaveSpeeds.UNIT)])
def test_cfl_over_the_live_list_is_bitwise_the_cfl_over_every_tile(threshold, batch_cfl, wave):
    """The CFL reduction over the live tiles, with the zero-state tiles taken
    by their materials and inflow columns, gives the maximum of every cell
    of every tile bit for bit: same energy grid, same dose."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg(tile_threshold=threshold, batch_cfl=batch_cfl)
    num = _native.numerics_struct(cfg)
    num.wave_speeds = 0 if wave is WaveSpeeds.EXACT else 1
    h, _ = prepare_batch(ph, beamlets, cfg, boxes)
    d0, s0 = h.march(num)
    h.set_cfl_all_tiles(True)
    d1, s1 = h.march(num)
    h.set_cfl_all_tiles(False)
    ev, full = h.tile_stats()
    say("CFL forms:", threshold, batch_cfl, wave.name, "steps", s0.tolist(),
        "bitwise", bool(np.array_equal(d0, d1)))
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1)


def test_graph_steps_are_bitwise_the_launched_steps():
    """Plain steps replayed from captured HIP graphs (persistent stage
    kernels, the next tile set and CFL reduction inside the graph) give the
    march of separate launches bit for bit."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg()
    num = _native.numerics_struct(cfg)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes)
    d0, s0 = h.march(num)
    h.set_graphs(True)
    d1, s1 = h.march(num)
    n_graph, err = h.graph_info()
    # with kernel timing on, every step is launched one by one
    h.set_profiling(True)
    d2, s2 = h.march(num)
    n_timed, _ = h.graph_info()
    h.set_profiling(False)
    h.set_graphs(False)
    say("graph steps", n_graph, "of", int(s1.max()), "capture error", repr(err),
        "graph steps while timed", n_timed)
    assert err == "" and n_graph > 0.9 * int(s1.max()) and n_timed == 0
    assert np.array_equal(s0, s1) and np.array_equal(s0, s2)
    assert np.array_equal(d0, d1) and np.array_equal(d0, d2)


@pytest.mark.parametrize("shape", range(1, len(_native.batch_tile_shapes())))
def test_every_tile_shape_gives_the_same_march(shape):
    """The kernels of each compiled tile shape evaluate the same arithmetic on
    a different tile set: the dose agrees with the default shape to the
    sub-threshold content the sets differ by, and against the FP64 march
    within the acceptance bounds."""
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg()
    num = _native.numerics_struct(cfg)
    h0, _ = prepare_batch(ph, beamlets, cfg, boxes)
    d0, s0 = h0.march(num)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes, tile_shape=shape)
    h.reset_stats()
    d1, s1 = h.march(num)
    ev, full = h.tile_stats()
    err = metrics.max_abs_difference_over_max(d1, d0)
    md = metrics.maximum_dose_relative_difference(d1, d0)
    say("tile shape", h.tile_shape, "tiles", ev, "of", full, "max|dD|/maxD vs 16^3", err,
        "max-dose", md, "steps", s1.tolist(), s0.tolist())
    assert h.tile_shape == _native.batch_tile_shapes()[shape]
    assert err < 2e-3 and abs(md) < 1.5e-3
    # the per-tile and persistent launches of this shape agree bit for bit
    h.set_persistent(True)
    d2, s2 = h.march(num)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)


def ragged_workload():
    """Boxes of 27 x 30 x depth cells (no axis a multiple of any tile
    dimension, depth not rounded), one clamped into a corner of the phantom."""
    ph, beamlets, _ = small_workload()
    beamlets = beamlets + [Beamlet(energy=50.0, sigma=(0.002, 0.002), center=(-0.023, -0.023))]
    boxes = beamlet_boxes(ph, beamlets, lateral=(27, 30), z_multiple=1)
    return ph, beamlets, boxes


@pytest.mark.parametrize("shape", range(len(_native.batch_tile_shapes())))
def test_ragged_boxes_equal_single_box_marches(shape):
    """Every tile evaluated (threshold 0) and the box CFL: each beamlet of the
    batch against its box marched alone.  With the 16 x 16 y-z tiles of the
    single-box kernels (shape 0) the doses are bitwise equal; other shapes
    place the y-z tile borders elsewhere, where the halo lanes form the lower
    WENO state from the mirrored stencil (the same value to rounding).
    Observed: 8 x 16 x 16 bitwise as well (x tile borders form the upper
    state with the same arithmetic as the interior); 16 x 8 x 32 and
    32 x 8 x 32 the same steps and <= 6.4e-4 of the maximum."""
    ph, beamlets, boxes = ragged_workload()
    assert boxes[-1][:2] == (0, 0)
    assert all(b[2] == 27 and b[3] == 30 for b in boxes)
    cfg = _Cfg(tile_threshold=0.0, batch_cfl="box")
    num = _native.numerics_struct(cfg)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes, tile_shape=shape)
    say("tile shape", h.tile_shape, "boxes", boxes)
    _, steps = h.march(num, want_dose=False)
    ctx = _native.Context(0)
    for b, (box, bl) in enumerate(zip(boxes, beamlets)):
        ctx.set_problem(*box_problem(ph, box, bl))
        E, _ = ctx.march(num)
        single = ctx.get_dose()
        got = h.beamlet_dose(b)
        err = metrics.max_abs_difference_over_max(got, single)
        say("  beamlet", b, "box", box, "steps", int(steps[b]), len(E) - 1, "max|dD|/maxD", err)
        assert got.shape == single.shape
        if shape == 0:
            assert steps[b] == len(E) - 1 and np.array_equal(got, single)
        else:
            assert abs(int(steps[b]) - (len(E) - 1)) <= 2 and err < 2e-3


def test_tile_shape_must_precede_the_problem():
    ph, beamlets, boxes = small_workload()
    h, _ = prepare_batch(ph, beamlets, _Cfg(), boxes)
    with pytest.raises(ValueError, match="before"):
        h.set_tile_shape(1)
    with pytest.raises(ValueError, match="out of range"):
        _native.Batch(0).set_tile_shape(99)


def test_autotune_picks_a_timed_candidate():
    from proton_solver.autotune import MODES, PREFER_WITHIN, autotune

    ph, beamlets, boxes = small_workload()
    choice = autotune(ph, beamlets, _Cfg(), sample=3)
    say("autotune", choice.shape, choice.mode, choice.seconds, choice.table)
    shapes = _native.batch_tile_shapes()
    assert len(choice.table) == len(shapes) + len(MODES) - 1
    assert choice.shape == shapes[choice.shape_index]
    # the fastest candidate, or graph replay within PREFER_WITHIN of it
    best = min(r["seconds"] for r in choice.table)
    assert choice.seconds <= best * (1.0 + PREFER_WITHIN) + 1e-3
    assert any(r["mode"] == choice.mode and tuple(r["shape"]) == choice.shape
               for r in choice.table)
    assert choice.mode == "device_graphs" or abs(choice.seconds - best) < 1e-3
