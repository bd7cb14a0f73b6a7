"""The lateral water / cortical-bone interface workload with its tile-skip
network, on an MI355X.

160 x 160 x 480 cells at 0.5 mm, cortical bone (1.85 g/cm^3) filling x > 0,
one 200 MeV beam (sigma 3 mm) entering 2 mm from the interface on the water
side, Wentzel + nuclear, IF-SSPRK3, FP32 bulk with tile skipping.  The beam
is solved by the batched march as one beamlet whose box is the whole grid,
with the network trained on full-domain runs of this workload at seeded beam
offsets (proton_solver.tileskip --workload interface; weights in
tileskip_weights_interface.npy).  Checked against the plain FP64 march on the
same grid (bitwise the CPU oracle): maximum-dose error <= 0.5 % and
gamma(1 %/1 mm) >= 99.5 % with and without the network, and the network
evaluates fewer tiles than the dilated mask alone.
"""

import numpy as np
import pytest

from proton_solver import _native, metrics, tileskip
from proton_solver.batch import (batch_workload_config, full_problem, interface_workload,
                                 prepare_batch)
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def interface_runs():
    cfg = batch_workload_config()
    num = _native.numerics_struct(cfg)
    phantom, beamlets, boxes = interface_workload(0.0)
    handle, _ = prepare_batch(phantom, beamlets, cfg, boxes=boxes)
    runs = {}
    model = tileskip.load_model(0, tileskip.INTERFACE_WEIGHTS_FILE)
    for label, m in (("mask", None), ("network", model)):
        tileskip.enable(handle, m, tileskip.SkipOptions(judge_every=256))
        handle.reset_stats()
        _, steps = handle.march(num, want_dose=False)
        tiles, full = handle.tile_stats()
        runs[label] = (handle.beamlet_dose(0), int(steps[0]), tiles, full)
        say(f"interface, {label}: {int(steps[0])} steps, tiles {tiles / full:.4f} of the grid's")
    tileskip.enable(handle, None)
    ctx = _native.Context(0)
    ctx.set_problem(*full_problem(phantom, beamlets[0]))
    ctx.set_state(None)
    E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
    say(f"interface, plain FP64: {len(E64) - 1} steps")
    return runs, ctx.get_dose(), phantom


@pytest.mark.parametrize("label", ["mask", "network"])
def test_interface_against_plain_fp64(interface_runs, label):
    runs, ref, phantom = interface_runs
    dose = runs[label][0]
    err = metrics.maximum_dose_relative_difference(dose, ref)
    gamma = metrics.gamma_pass_rate(dose, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0)
    say(f"interface, {label}: max-dose error {err:+.2e}, gamma {gamma:.3f} %")
    assert abs(err) <= 5e-3 and gamma >= 99.5


def test_interface_network_evaluates_fewer_tiles(interface_runs):
    runs, _, _ = interface_runs
    assert runs["network"][2] < 0.85 * runs["mask"][2]
