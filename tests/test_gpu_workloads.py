"""The single-beamlet benchmark workloads at full size, FP32 bulk against FP64.

Both marches run on the device: the FP64 march is bitwise the oracle (the
parity tests at oracle-sized grids), so it stands in for the oracle at
sizes the CPU restatement would need minutes for.  Acceptance, as stated for
the mixed-precision march: maximum-dose error <= 0.5 % and
gamma(1 %/1 mm, 10 % cutoff) >= 99.5 % against FP64, for both float
closures.

  * water: 150 MeV, sigma 4 mm, 120 x 120 x 360 cells of 0.5 mm, Wentzel
    scattering and nuclear removal, max step 0.25 MeV;
  * interface: cortical bone (1.85 g/cm^3) filling x > 0, 200 MeV, sigma
    3 mm, centred 2 mm on the water side (x = -2 mm), 160 x 160 x 480 cells
    of 0.5 mm, the same physics.  The water-side range (about 26 cm) exceeds
    the 24 cm depth, so part of the beam leaves through the z-max face.

Observed on an MI355X (float / double closure): water 919 steps in FP64 and
FP32, maximum-dose error -3.1e-4 / 5.0e-4, gamma 100 %; interface 1437 FP64
steps against 1440, maximum-dose error 6.2e-5 / 1.3e-3, largest local
difference 1.5e-2 / 3.0e-2 of the maximum (where the grids' three extra steps
fall), gamma 100 %.
"""

import sys
import time

import numpy as np
import pytest

from proton_solver import (Beam, Box, DeviceOptions, EnergyScheme, Geometry, Material, Numerics,
                           Physics, Scattering, SolverConfig, SpatialScheme, WaveSpeeds, _native,
                           metrics)
from proton_solver.numerics.boundary import BeamInflow
from proton_solver.numerics.grid import Grid
from proton_solver.physics.materials import MaterialMap

pytestmark = pytest.mark.gpu

MAX_DOSE_TOL = 5e-3
GAMMA_MIN = 99.5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _config(size, resolution, objects, energy, sigma, center):
    return SolverConfig(
        geometry=Geometry(size=size, resolution=resolution, objects=list(objects)),
        beam=Beam(n_protons=1e6, energy=energy, energy_spread=1.0, sigma=(sigma, sigma),
                  center=center, angular_spread=0.0, cutoff_sigma=5.0),
        physics=Physics(scattering=Scattering.WENTZEL, nuclear=True),
        numerics=Numerics(spatial_scheme=SpatialScheme.WENO_Z,
                          energy_scheme=EnergyScheme.IF_SSPRK3, wave_speeds=WaveSpeeds.EXACT,
                          max_energy_step=0.25, cfl=0.9, energy_min=1.0),
        device=DeviceOptions(),
    )


WORKLOADS = {
    "water_150MeV_120x120x360": _config((0.06, 0.06, 0.18), (120, 120, 360), (), 150.0, 0.004,
                                        (0.0, 0.0)),
    "bone_interface_200MeV_160x160x480": _config(
        (0.08, 0.08, 0.24), (160, 160, 480),
        (Box(0.0, 0.04, -0.04, 0.04, 0.0, 0.24, Material.BONE_COMPACT),), 200.0, 0.003,
        (-0.002, 0.0)),
}


@pytest.fixture(scope="module")
def fp64_runs():
    return {}


@pytest.mark.parametrize("closure", ["float_closure", "f64_closure"])
@pytest.mark.parametrize("workload", list(WORKLOADS))
def test_workload_fp32_bulk_against_fp64(workload, closure, fp64_runs):
    cfg = WORKLOADS[workload]
    grid = Grid.from_geometry(cfg.geometry, np)
    mm = MaterialMap.from_geometry(cfg.geometry, grid, np)
    inflow = BeamInflow.from_beam(cfg.beam, grid, np)
    say(workload, closure, "grid", grid.shape)
    ctx = _native.Context(0)
    ctx.set_problem(grid, mm, inflow, cfg.beam)
    if workload not in fp64_runs:
        ctx.set_state(None)
        t0 = time.perf_counter()
        E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
        fp64_runs[workload] = (ctx.get_dose(), len(E64) - 1, time.perf_counter() - t0)
        say("  FP64", len(E64) - 1, "steps", f"{fp64_runs[workload][2]:.2f} s")
    ref, steps64, _ = fp64_runs[workload]
    precision = _native.FP32_BULK if closure == "float_closure" else _native.FP32_BULK_F64_CLOSURE
    ctx.set_state(None)
    t0 = time.perf_counter()
    E, _ = ctx.march(_native.numerics_struct(cfg, precision))
    sec = time.perf_counter() - t0
    dose = ctx.get_dose()
    md = metrics.maximum_dose_relative_difference(dose, ref)
    mx = metrics.max_abs_difference_over_max(dose, ref)
    gam = metrics.gamma_pass_rate(dose, ref, (grid.dx, grid.dy, grid.dz), 1.0, 1.0, 10.0,
                                  device=0)
    say(f"  FP32 bulk ({closure}) {len(E) - 1} steps {sec:.2f} s  max-dose err {md:.3e}  "
        f"max|dD|/max {mx:.3e}  gamma(1%/1mm) {gam:.3f} %")
    assert np.all(np.isfinite(dose)) and np.all(dose >= 0)
    assert abs(md) <= MAX_DOSE_TOL
    assert gam >= GAMMA_MIN


# Beamlets of the combined workload (1024 beamlets, 256 x 256 x 200 phantom,
# five materials) whose box and a +-48 mm footprint around the entry point
# hold no air over the box depth (131 of the 1024; the first six here, 73 to
# 147 MeV).
COMBINED_AIR_FREE = (1, 3, 19, 21, 23, 29)


def test_combined_workload_air_free_beamlets():
    """Six beamlets of the combined five-material workload in one batch, with
    the double closure (FP32_BULK_F64_CLOSURE), the vacuum-aware CFL, the
    tile-skip network and judge, and graph steps, against the FP64
    moving-domain march with the vacuum-aware CFL (the reference's own march
    for phantoms holding air, energy.py:1262-1494) of each beamlet on the
    whole phantom.  The plain FP64 march is not the comparison here: any air
    cell of the phantom bounds its step at 1/934 of water's.  Beamlets near
    air are left out for the same reason (their steps are air-bound while the
    beam passes it).  Observed on an MI355X: the same energy steps in both
    (309 - 604), maximum-dose error -1.7e-4 .. 4.9e-4, gamma 100 %."""
    from proton_solver import tileskip
    from proton_solver.batch import (air_free_beamlets, batch_workload_config, beamlet_boxes,
                                     combined_workload, full_problem, prepare_batch)

    phantom, beamlets = combined_workload()
    air = phantom.materials.index(Material.AIR)
    assert air_free_beamlets(phantom, beamlets, beamlet_boxes(phantom, beamlets))[:6] == \
        list(COMBINED_AIR_FREE)
    chosen = [beamlets[i] for i in COMBINED_AIR_FREE]
    cfg = batch_workload_config(vacuum_aware_cfl=True)
    handle, boxes = prepare_batch(phantom, chosen, cfg)
    for (x0, y0, nx, ny, nz), bl in zip(boxes, chosen):
        ix = int(np.floor(bl.center[0] / 0.001)) + 128
        iy = int(np.floor(bl.center[1] / 0.001)) + 128
        foot = phantom.ids[max(ix - 48, 0):ix + 48, max(iy - 48, 0):iy + 48, :nz]
        assert not (foot == air).any()
    model = tileskip.load_model(0) if tileskip.WEIGHTS_FILE.exists() else None
    tileskip.enable(handle, model, tileskip.SkipOptions())
    handle.set_graphs(True)
    say("combined workload: beamlets", COMBINED_AIR_FREE, "network", model is not None)
    t0 = time.perf_counter()
    _, steps = handle.march(_native.numerics_struct(cfg, _native.FP32_BULK_F64_CLOSURE),
                            want_dose=False)
    say(f"  batch {time.perf_counter() - t0:.2f} s, steps {list(steps)}")
    handle.set_graphs(False)
    ctx = _native.Context(0)
    num64 = _native.numerics_struct(cfg, _native.FP64)
    for b, bl in enumerate(chosen):
        x0, y0, nx, ny, nz = boxes[b]
        got = handle.beamlet_dose(b)
        ctx.set_problem(*full_problem(phantom, bl))
        ctx.set_state(None)
        t0 = time.perf_counter()
        E, _, mboxes = ctx.march_moving(num64, 9, 5.0)
        sec = time.perf_counter() - t0
        full = ctx.get_dose()
        ref = full[x0:x0 + nx, y0:y0 + ny, :nz]
        outside = 1.0 - ref.sum() / full.sum()
        md = got.max() / full.max() - 1.0
        gam = metrics.gamma_pass_rate(got, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0)
        say(f"  beamlet {COMBINED_AIR_FREE[b]} ({bl.energy:.1f} MeV): batch {steps[b]} steps, "
            f"FP64 moving {len(E) - 1} steps {sec:.2f} s, dose outside the box {outside:.1e}, "
            f"max-dose err {md:.3e}, gamma(1%/1mm) {gam:.3f} %")
        assert np.all(np.isfinite(got)) and np.all(got >= 0)
        assert outside < 1e-3
        assert abs(md) <= MAX_DOSE_TOL
        assert gam >= GAMMA_MIN
