"""Gamma index on the device (m1_gamma_pass_count) against the host search.

proton_solver.metrics.gamma_pass_rate is pinned on the CPU
(tests/test_api.py: identical and shifted doses, a point checked by hand).
The device kernel evaluates the same points and offsets with the same double
arithmetic, so its pass count must equal the host's exactly -- on smooth,
noisy, shifted and scaled doses, several criteria, grids with a one-cell
axis, and a beamlet dose of the batch against its FP64 march.
"""

import sys
import time

import numpy as np
import pytest

from proton_solver import metrics

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bragg_like(shape, seed, shift=(0.0, 0.0, 0.0), scale=1.0, noise=0.0):
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    cx, cy = (shape[0] - 1) / 2 + shift[0], (shape[1] - 1) / 2 + shift[1]
    zp = 0.7 * shape[2] + shift[2]
    lat = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * (0.15 * shape[0]) ** 2))
    depth = 0.3 + 0.7 * np.exp(-((z - zp) ** 2) / 8.0) * (z < zp + 6)
    d = scale * lat * depth
    if noise:
        d = d * (1.0 + noise * rng.standard_normal(shape))
    return np.maximum(d, 0.0)


# (shape, evaluated-dose change, (dose %, distance mm)); host pass rates
# 47.3 %, 93.6 %, 99.95 %, 100 % and 0 %
CASES = [
    ((24, 20, 40), dict(shift=(1.5, 0.0, 1.2)), (1.0, 1.0)),
    ((24, 20, 40), dict(noise=0.05, shift=(0.0, 1.6, -1.8)), (2.0, 2.0)),
    ((24, 20, 40), dict(shift=(0.3, 0.0, 0.4)), (1.0, 1.0)),
    ((24, 20, 40), dict(noise=0.01), (1.0, 1.0)),
    ((1, 16, 30), dict(shift=(0.0, 2.5, 3.0)), (1.0, 1.0)),
]


@pytest.mark.parametrize("shape,kw,crit", CASES)
def test_device_gamma_equals_host(shape, kw, crit):
    ref = bragg_like(shape, 0)
    ev = bragg_like(shape, 1, **kw)
    sp = (1e-3, 1.2e-3, 0.8e-3)
    host = metrics.gamma_pass_rate(ev, ref, sp, *crit)
    dev = metrics.gamma_pass_rate(ev, ref, sp, *crit, device=0)
    say(shape, kw, crit, "host", host, "device", dev)
    assert dev == host
    assert 0.0 <= host <= 100.0


def test_device_gamma_on_a_batch_beamlet():
    from proton_solver import _native
    from proton_solver.batch import batch_workload, batch_workload_config, full_problem, prepare_batch

    phantom, beamlets = batch_workload(n_beamlets=4)
    cfg = batch_workload_config()
    handle, boxes = prepare_batch(phantom, beamlets[:1], cfg)
    handle.march(_native.numerics_struct(cfg), want_dose=False)
    x0, y0, nx, ny, nz = boxes[0]
    got = handle.beamlet_dose(0)
    ctx = _native.Context(0)
    grid, mm, inflow, beam = full_problem(phantom, beamlets[0])
    ctx.set_problem(grid, mm, inflow, beam)
    ctx.march(_native.numerics_struct(cfg, _native.FP64))
    ref = ctx.get_dose()[x0:x0 + nx, y0:y0 + ny, :nz]
    t0 = time.perf_counter()
    host = metrics.gamma_pass_rate(got, ref, phantom.spacing, 1.0, 1.0, 10.0)
    t1 = time.perf_counter()
    dev = metrics.gamma_pass_rate(got, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0)
    t2 = time.perf_counter()
    say("beamlet", beamlets[0].energy, "box", (nx, ny, nz), "gamma host", host,
        f"{t1 - t0:.2f} s", "device", dev, f"{t2 - t1:.3f} s")
    assert dev == host
    assert host >= 99.5


def test_device_gamma_argument_checks():
    from proton_solver import _native

    d = np.zeros((2, 2, 2))
    with pytest.raises(ValueError, match="outside"):
        _native.gamma_pass_count(0, d, d, np.array([[0, 0, 5]]), np.zeros((1, 3)),
                                 np.zeros(1), 1.0)
    with pytest.raises(ValueError, match="sorted"):
        _native.gamma_pass_count(0, d, d, np.array([[0, 0, 0]]), np.zeros((2, 3)),
                                 np.array([1.0, 0.5]), 1.0)
