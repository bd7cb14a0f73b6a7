"""Device-side energy stepping of the batched march (m1_batch_set_device_steps),
on an MI355X, through the C ABI.

Checked:
  * the device's step physics -- S(E) (stopping_power.py:141-228), the
    transport and removal coefficients (scattering.py:65-116,
    nuclear.py:94-138, linear and PCHIP) and the Gauss-Legendre optical
    depths (scattering.py:267-306, nuclear.py:202-233) -- equals the oracle's
    bit for bit at a few thousand energies per material (these use only
    sums, products, quotients and square roots), and an energy outside a
    table raises the reference's ValueError;
  * a device-stepped march takes the host-stepped march's energy steps and
    its dose agrees to float rounding (the inflow spectrum and the
    integrating factors take their exponentials from the device library);
  * the estimated-grid launch, the persistent launch and graph replay of
    device-stepped steps are the same march bit for bit;
  * with the tile-skip network (beamlets judged on the batch's common
    schedule): a network that never skips is the plain device-stepped march
    bit for bit; an always-skipping network judged at 1e-12 is rolled back
    and a beamlet sent back to its start equals the plain march bit for bit.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as OR
from proton_solver import _native, metrics
from proton_solver import tileskip as TS
from proton_solver.batch import (
    FIVE_MATERIALS, Beamlet, Phantom, batch_workload_config, prepare_batch,
)
from test_gpu_batch import _Cfg, small_workload
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _five_material_batch():
    """A small phantom holding the five tabulated materials (slabs along x)."""
    shape = (40, 40, 48)
    ids = np.zeros(shape, dtype=np.int32)
    for m in range(len(FIVE_MATERIALS)):
        ids[8 * m:8 * (m + 1)] = m
    ph = Phantom(ids=ids, materials=FIVE_MATERIALS, spacing=(0.001,) * 3)
    beamlets = [Beamlet(energy=40.0, sigma=(0.002, 0.002), center=(0.0, 0.0))]
    handle, _ = prepare_batch(ph, beamlets, _Cfg(), boxes=[(4, 4, 32, 32, 48)])
    return handle


def _oracle_stopping_power(mat, e):
    out = C.c_double()
    rc = OR.lib().or_stopping_power(C.byref(mat), C.c_double(e), C.byref(out))
    assert rc == 0
    return out.value


@pytest.mark.parametrize("interp", ["LINEAR", "PCHIP"])
def test_device_step_physics_is_the_oracle_bitwise(interp):
    handle = _five_material_batch()
    rng = np.random.default_rng(7)
    energies = np.concatenate([rng.uniform(1.0, 260.0, 1500), rng.uniform(1e-4, 1.0, 200),
                               [1e-4, 300.0, 299.9999999, 150.0, 0.5]])
    code = {"LINEAR": 0, "PCHIP": 1}[interp]
    for m, material in enumerate(FIVE_MATERIALS):
        mat, keep = OR.material_struct(material)
        sp = handle.physics_probe(m, 0, energies)
        ref = np.array([_oracle_stopping_power(mat, float(e)) for e in energies])
        assert np.array_equal(sp, ref), (material, interp, "S")
        for kind, tab in ((1, mat.tr), (2, mat.nu)):
            Et = np.ctypeslib.as_array(tab.E, shape=(tab.n,)).copy()
            vt = np.ctypeslib.as_array(tab.v, shape=(tab.n,)).copy()
            got = handle.physics_probe(m, kind, energies, scattering=2, interpolation=code)
            ref = np.array([OR.table_eval(Et, vt, float(e), interp) for e in energies])
            assert np.array_equal(got, ref), (material, interp, kind)
        # optical depths over intervals inside the tables
        hi = energies[(energies > 2.0)][:600]
        for dE in (0.25, 0.0173, 1.7):
            got = handle.physics_probe(m, 3, hi, dE=dE, scattering=2, interpolation=code)
            ref = np.array([OR.scattering_depth(mat, "WENTZEL", float(e), dE, interp) for e in hi])
            assert np.array_equal(got, ref), (material, interp, "tau_s", dE)
            got = handle.physics_probe(m, 4, hi, dE=dE, interpolation=code)
            ref = np.array([OR.nuclear_depth(mat, float(e), dE, interp) for e in hi])
            assert np.array_equal(got, ref), (material, interp, "tau_n", dE)
        got = handle.physics_probe(m, 3, hi[:50], dE=0.3, scattering=1, interpolation=code)
        ref = np.array([OR.scattering_depth(mat, "ROSSI", float(e), 0.3, interp) for e in hi[:50]])
        assert np.array_equal(got, ref), (material, "Rossi")
        del keep
    say("device step physics bitwise the oracle,", interp, len(energies), "energies x 5 materials")


def test_device_step_physics_range_error():
    handle = _five_material_batch()
    with pytest.raises(ValueError, match="outside stopping-power table range"):
        handle.physics_probe(1, 0, np.array([150.0, 400.0]))
    with pytest.raises(ValueError, match="outside nuclear table range"):
        handle.physics_probe(1, 2, np.array([-3.0]))
    with pytest.raises(ValueError, match="kind"):
        handle.physics_probe(1, 7, np.array([1.0]))


@pytest.fixture(scope="module")
def host_stepped():
    ph, beamlets, boxes = small_workload()
    cfg = _Cfg(tile_threshold=1e-6)
    num = _native.numerics_struct(cfg)
    h, _ = prepare_batch(ph, beamlets, cfg, boxes)
    dose, steps = h.march(num)
    per = [h.beamlet_dose(b) for b in range(len(beamlets))]
    return h, num, dose, steps, per


def test_device_stepped_march_equals_host_stepped(host_stepped):
    h, num, dose0, steps0, per0 = host_stepped
    h.reset_stats()
    h.set_device_steps(True)
    try:
        dose, steps = h.march(num)
        ev, full = h.tile_stats()
        per = [h.beamlet_dose(b) for b in range(len(per0))]
    finally:
        h.set_device_steps(False)
    h.reset_stats()
    h.march(num)
    ev0, full0 = h.tile_stats()
    diff = max(float(np.abs(a - b).max() / b.max()) for a, b in zip(per, per0))
    say("device steps: steps", steps.tolist(), "host", steps0.tolist(), "bitwise",
        bool(np.array_equal(dose, dose0)), "max|dD|/max", diff, "tiles", ev, ev0, full, full0)
    assert np.array_equal(steps, steps0)
    assert diff <= 1e-5
    assert full == full0 and abs(ev - ev0) <= 0.01 * ev0


def test_device_stepped_launch_forms_are_bitwise_the_same_march(host_stepped):
    h, num, *_ = host_stepped
    h.set_device_steps(True)
    try:
        d_est, s_est = h.march(num)
        h.set_persistent(True)
        d_per, s_per = h.march(num)
        h.set_persistent(False)
        h.set_graphs(True)
        d_gr, s_gr = h.march(num)
        n_graph, err = h.graph_info()
        h.set_graphs(False)
        h.set_profiling(True)
        d_pr, s_pr = h.march(num)
        stats = h.kernel_stats()
        h.set_profiling(False)
    finally:
        h.set_device_steps(False)
        h.set_graphs(False)
        h.set_persistent(False)
        h.set_profiling(False)
    say("device-stepped forms: graph steps", n_graph, "error", repr(err), "step kernels",
        stats.get("step_scalars"), stats.get("step_advance"))
    assert err == "" and n_graph > 0.9 * int(s_gr.max())
    for d, s in ((d_per, s_per), (d_gr, s_gr), (d_pr, s_pr)):
        assert np.array_equal(s, s_est)
        assert np.array_equal(d, d_est)
    assert stats["step_scalars"][0] >= int(s_est.max())


def test_device_stepped_march_against_fp64(host_stepped):
    """Device stepping keeps the batch's acceptance against the plain FP64
    march of each beamlet on the whole phantom."""
    from proton_solver.batch import full_problem

    h, num, *_ = host_stepped
    ph, beamlets, boxes = small_workload()
    h.set_device_steps(True)
    try:
        dose, steps = h.march(num)
    finally:
        h.set_device_steps(False)
    from proton_solver import Precision

    ctx = _native.Context(0)
    cfg64 = _Cfg(precision=Precision.FP64)
    num64 = _native.numerics_struct(cfg64)
    b = 1
    ctx.set_problem(*full_problem(ph, beamlets[b]))
    ctx.march(num64)
    ref = ctx.get_dose()
    x0, y0, nx, ny, nz = boxes[b]
    got = np.zeros_like(ref)
    got[x0:x0 + nx, y0:y0 + ny, :nz] = h.beamlet_dose(b)
    md = metrics.maximum_dose_relative_difference(got, ref)
    g = metrics.gamma_pass_rate(got, ref, ph.spacing, 1.0, 1.0, 10.0)
    say("device-stepped beamlet", b, "vs FP64: max-dose", md, "gamma", g)
    assert abs(md) <= 5e-3 and g >= 99.5


def _workload(phantom_seed, beamlet_seed, n):
    from proton_solver.batch import ellipsoid_phantom, random_beamlets, random_ellipsoids

    shape, spacing = (128, 128, 128), (0.001,) * 3
    size = tuple(a * b for a, b in zip(shape, spacing))
    ph = ellipsoid_phantom(shape, spacing,
                           random_ellipsoids(np.random.default_rng(phantom_seed), 4, size))
    beamlets = random_beamlets(np.random.default_rng(beamlet_seed), n, energy_range=(60.0, 110.0))
    return ph, beamlets


def _constant_model(logit):
    p = np.zeros(_native.load_library().m1_mlp_param_count(), dtype=np.float32)
    p[-1] = logit
    return _native.Mlp(0, p)


@pytest.fixture(scope="module")
def net_target():
    cfg = batch_workload_config()
    ph, beamlets = _workload(2, 20, 8)
    handle, _ = prepare_batch(ph, beamlets, cfg)
    num = _native.numerics_struct(cfg)
    handle.set_device_steps(True)
    dose, steps = handle.march(num)
    per = [handle.beamlet_dose(b) for b in range(len(beamlets))]
    yield handle, num, dose, steps, per
    handle.set_device_steps(False)


def _march(handle, num, model, options):
    TS.enable(handle, model, options)
    try:
        handle.reset_stats()
        dose, steps = handle.march(num)
        tiles, _ = handle.tile_stats()
        judged, rejections, back = handle.skip_stats()
        per = [handle.beamlet_dose(b) for b in range(len(handle.boxes))]
    finally:
        TS.enable(handle, None)
    return dose, steps, tiles, judged, rejections, back, per


def test_device_stepped_never_skipping_network_is_the_plain_march(net_target):
    handle, num, dose0, steps0, per0 = net_target
    dose, steps, tiles, judged, rejections, back, _ = _march(
        handle, num, _constant_model(-30.0), TS.SkipOptions(judge_every=4))
    say("device steps, never-skip network: judged", judged, "rejections", rejections)
    assert judged > 0 and rejections == 0 and (back == -1).all()
    assert np.array_equal(steps, steps0)
    assert np.array_equal(dose, dose0)


def test_device_stepped_judge_rolls_back(net_target):
    handle, num, dose0, steps0, per0 = net_target
    dose, steps, tiles, judged, rejections, back, per = _march(
        handle, num, _constant_model(30.0),
        TS.SkipOptions(judge_every=8, judge_tol=1e-12, guard=4))
    say("device steps, always-skip network at 1e-12: judged", judged, "rejections", rejections,
        "rolled back to", back.tolist())
    assert rejections > 0 and (back >= 0).sum() == rejections
    for b, (got, ref) in enumerate(zip(per, per0)):
        if back[b] == 0:
            assert np.array_equal(got, ref), b
            assert steps[b] == steps0[b]
        assert np.abs(got - ref).max() <= 1e-3 * ref.max(), b
    # graph replay of the plain steps, judged steps launched: the same march
    handle.set_graphs(True)
    try:
        d_g, s_g, _, judged_g, rej_g, back_g, _ = _march(
            handle, num, _constant_model(30.0),
            TS.SkipOptions(judge_every=8, judge_tol=1e-12, guard=4))
        n_graph, err = handle.graph_info()
    finally:
        handle.set_graphs(False)
    say("with graphs: graph steps", n_graph, "error", repr(err))
    assert err == "" and n_graph > 0
    assert np.array_equal(s_g, steps) and np.array_equal(d_g, dose)
    assert judged_g == judged and rej_g == rejections and np.array_equal(back_g, back)
