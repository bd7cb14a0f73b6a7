"""Quasi-steady air in the batched march (m1_batch_set_quasi_steady), on an
MI355X, through the C ABI.

Air's stopping power is 1/934 of water's, so the reference's march on a
phantom holding air takes its CFL step from the air cells: the plain FP64
march of a 60 MeV beam through a 10 mm air pocket (48 x 48 x 64 cells at
1 mm) takes ~43,000 energy steps where the water alone needs ~250.  With the
air quasi-steady the batch takes water-sized steps, holds the air cells
through each stage and relaxes them towards div F = 0 after it.  Checked
against that plain FP64 march (bitwise the CPU oracle): maximum dose within
0.5 %, gamma(1 %/1 mm) >= 99.5 %, with a 0.1 MeV maximum step (the plain
march's own energy grid is ~1e-3 MeV here; at 0.25 MeV the water alone
differs from a converged energy grid by ~0.5 % in maximum dose).  The same
with the sweeps stopping per beamlet at a tolerance
(m1_batch_set_quasi_steady_tolerance), the convergence machinery bitwise
neutral when no box converges, a beamlet of the combined five-material
workload with air beside its axis against the plain FP64 march of its box,
and the argument checks.
"""

import dataclasses

import numpy as np
import pytest

from proton_solver import Material, _native, metrics
from proton_solver.autotune import set_mode
from proton_solver.batch import (
    Beamlet, Ellipsoid, batch_workload_config, beamlet_boxes, box_problem, combined_workload,
    ellipsoid_phantom, full_problem, prepare_batch,
)
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _problem(max_step=0.1):
    ph = ellipsoid_phantom((48, 48, 64), (0.001,) * 3,
                           [Ellipsoid((0.0, 0.0, 0.015), (0.008, 0.008, 0.005), Material.AIR)])
    beamlets = [Beamlet(energy=60.0, sigma=(0.003, 0.003), center=(0.0, 0.0))]
    cfg = batch_workload_config(tile_threshold=1e-6)
    cfg = dataclasses.replace(cfg, numerics=dataclasses.replace(cfg.numerics,
                                                                max_energy_step=max_step))
    return ph, beamlets, cfg


@pytest.fixture(scope="module")
def plain_fp64():
    ph, beamlets, cfg = _problem()
    ctx = _native.Context(0)
    ctx.set_problem(*full_problem(ph, beamlets[0]))
    ctx.set_state(None)
    say("plain FP64 march through the air pocket (air-bound steps)")
    E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
    return len(E64) - 1, ctx.get_dose()


def _quasi_steady(sweeps, tolerance):
    ph, beamlets, cfg = _problem()
    air = ph.materials.index(Material.AIR)
    handle, _ = prepare_batch(ph, beamlets, cfg, boxes=[(0, 0, 48, 48, 64)], tile_shape=2)
    set_mode(handle, "device")
    handle.set_quasi_steady(1 << air, sweeps, 0.3)
    handle.set_quasi_steady_tolerance(tolerance)
    _, steps = handle.march(_native.numerics_struct(cfg), want_dose=False)
    got = handle.beamlet_dose(0)
    handle.close()
    return ph, int(steps[0]), got


@pytest.mark.parametrize("sweeps,tolerance", [(64, 0.0), (256, 1e-6)])
def test_quasi_steady_air_against_the_plain_fp64_march(plain_fp64, sweeps, tolerance):
    n64, ref = plain_fp64
    ph, steps, got = _quasi_steady(sweeps, tolerance)
    err = metrics.maximum_dose_relative_difference(got, ref)
    diff = metrics.max_abs_difference_over_max(got, ref)
    gamma = metrics.gamma_pass_rate(got, ref, ph.spacing, 1.0, 1.0, 10.0, device=0)
    say(f"quasi-steady air ({sweeps} sweeps, tolerance {tolerance:g}): {steps} steps against "
        f"{n64}; max-dose error {err:+.2e}, max |dD| / max {diff:.2e}, gamma {gamma:.3f} %")
    assert n64 > 50 * steps
    assert abs(err) <= 5e-3
    assert gamma >= 99.5
    # the dose before the pocket (z < 9 mm) is the water march's
    assert np.abs(got[:, :, :9] - ref[:, :, :9]).max() <= 2e-3 * ref.max()


def test_convergence_flags_keep_the_fixed_sweeps():
    """With a tolerance no change can fall below (1e-30: only boxes whose air
    cells stop changing exactly are done), the convergence machinery --
    per-box flags and changes, the copy of a done box into both scratch
    states -- leaves the march bitwise that of the fixed sweep count."""
    _, steps0, fixed = _quasi_steady(32, 0.0)
    _, steps1, flagged = _quasi_steady(32, 1e-30)
    say(f"fixed sweeps against flags at 1e-30: steps {steps0} / {steps1}, max |dD| "
        f"{float(np.abs(fixed - flagged).max()):.3e}")
    assert steps0 == steps1
    assert np.array_equal(fixed, flagged)


def test_combined_workload_beamlet_near_air_against_the_plain_fp64_march():
    """The lowest-energy beamlet of the combined workload with air within
    5 mm of its axis in its box (beamlet 503, 78 MeV, 12,220 air cells):
    batched with the air quasi-steady (at most 128 sweeps, tolerance 1e-5,
    0.1 MeV maximum step) against the plain FP64 march of its box, which
    takes air-bound steps."""
    phantom, beamlets = combined_workload()
    cfg = batch_workload_config()
    cfg = dataclasses.replace(cfg, numerics=dataclasses.replace(cfg.numerics,
                                                                max_energy_step=0.1))
    n = cfg.numerics
    boxes = beamlet_boxes(phantom, beamlets, n.energy_min, n.energy_window_sigma)
    air = phantom.materials.index(Material.AIR)
    px, py, _ = phantom.shape
    dx = phantom.spacing[0]

    def air_distance(b):
        x0, y0, nx, ny, nz = boxes[b]
        cells = np.argwhere(phantom.ids[x0:x0 + nx, y0:y0 + ny, :nz] == air)
        if not len(cells):
            return np.inf
        ix = np.floor((beamlets[b].center[0] + 0.5 * px * dx) / dx) - x0
        iy = np.floor((beamlets[b].center[1] + 0.5 * py * dx) / dx) - y0
        return float(np.hypot(cells[:, 0] - ix, cells[:, 1] - iy).min())

    near = [b for b in np.argsort([bl.energy for bl in beamlets]) if air_distance(b) < 5.0]
    b = int(near[0])
    handle, _ = prepare_batch(phantom, [beamlets[b]], cfg, boxes=[boxes[b]], tile_shape=2)
    set_mode(handle, "device")
    handle.set_quasi_steady(1 << air, 128, 0.3)
    handle.set_quasi_steady_tolerance(1e-5)
    _, steps = handle.march(_native.numerics_struct(cfg), want_dose=False)
    got = handle.beamlet_dose(0)
    handle.close()
    ctx = _native.Context(0)
    ctx.set_problem(*box_problem(phantom, boxes[b], beamlets[b]))
    ctx.set_state(None)
    say(f"plain FP64 march of combined beamlet {b} ({beamlets[b].energy:.1f} MeV) on its box")
    E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
    ref = ctx.get_dose()
    err = metrics.maximum_dose_relative_difference(got, ref)
    gamma = metrics.gamma_pass_rate(got, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0)
    say(f"combined beamlet {b}: {int(steps[0])} steps against {len(E64) - 1}; max-dose error "
        f"{err:+.2e}, max |dD| / max {metrics.max_abs_difference_over_max(got, ref):.2e}, "
        f"gamma {gamma:.3f} %")
    assert len(E64) - 1 > 50 * int(steps[0])
    assert abs(err) <= 5e-3
    assert gamma >= 99.5


def test_quasi_steady_argument_checks():
    ph, beamlets, cfg = _problem()
    air = ph.materials.index(Material.AIR)
    handle, _ = prepare_batch(ph, beamlets, cfg, boxes=[(0, 0, 48, 48, 64)])
    num = _native.numerics_struct(cfg)
    for bad in ((1 << 8, 8, 0.5), (1 << air, -1, 0.5), (1 << air, 8, 0.0), (1 << air, 8, 1.5)):
        with pytest.raises(ValueError):
            handle.set_quasi_steady(*bad)
    for bad in (-1e-6, 1.0, float("nan")):
        with pytest.raises(ValueError):
            handle.set_quasi_steady_tolerance(bad)
    handle.set_quasi_steady(1 << air, 8, 0.3)
    set_mode(handle, "per_tile")  # host-side stepping
    with pytest.raises(NotImplementedError, match="device-side stepping"):
        handle.march(num, want_dose=False)
    handle.set_quasi_steady(0, 0, 0.5)
