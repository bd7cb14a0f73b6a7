"""The tile-skip network inside the batched march, on an MI355X.

A 128 x 128 x 128 phantom at 1 mm (water with four seeded bone / lung
ellipsoids) and beamlets of 60-110 MeV on 64 x 64 boxes.  Checked:
  * recording yields labelled samples of the candidate tiles with finite
    features in their documented ranges;
  * a network trained on the device from samples of another phantom and
    beamlet set leaves out at least 1.5 times the held-out candidates the
    fixed 4-cell guard rule does, at no more than 1e-4 false skips, and in
    the march evaluates fewer tiles than both the reach rule alone and the
    guard rule while the dose moves by less than 0.1 % of its maximum;
  * a network that never skips reproduces the march without the network
    bitwise, judge and all (the judge's save / full step / restore / skip
    step sequence is exact);
  * the judge's roll-back: a network that always skips, judged with a
    tolerance of 1e-12, is rejected; a beamlet sent back to its start
    equals the march without the network bitwise, and every dose stays
    within 0.1 % of it;
  * argument checks of the entry points.
"""

import numpy as np
import pytest

from proton_solver import _native
from proton_solver import tileskip as TS
from proton_solver.batch import (
    batch_workload_config, ellipsoid_phantom, prepare_batch, random_beamlets, random_ellipsoids,
)
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _workload(phantom_seed, beamlet_seed, n):
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
def cfg():
    return batch_workload_config()


@pytest.fixture(scope="module")
def samples(cfg):
    """Samples of every candidate (guard 0) on a training phantom and
    beamlet set (other seeds than the target below)."""
    ph, beamlets = _workload(1, 10, 24)
    handle, _ = prepare_batch(ph, beamlets, cfg)
    x, y, near = TS.record_samples(handle, _native.numerics_struct(cfg), horizon=8, guard=0)
    say("recorded samples", len(y), "empty fraction", float(y.mean()), "near fraction",
        float(near.mean()))
    x4, y4, near4 = TS.record_samples(handle, _native.numerics_struct(cfg), horizon=8, guard=4)
    say("guard-4 samples", len(y4))
    assert not near4.any()  # the 4-cell guard admits no such candidate
    return x, y, near


@pytest.fixture(scope="module")
def target(cfg):
    """The test problem and its march without the network."""
    ph, beamlets = _workload(2, 20, 12)
    handle, boxes = prepare_batch(ph, beamlets, cfg)
    num = _native.numerics_struct(cfg)
    handle.reset_stats()
    dose, steps = handle.march(num)
    tiles, full = handle.tile_stats()
    per = [handle.beamlet_dose(b) for b in range(len(beamlets))]
    say("no network: tiles", tiles, "of", full)
    return handle, num, dose, steps, tiles, per


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


def test_recorded_samples(samples, cfg):
    x, y, near = samples
    assert len(y) > 1000 and x.shape == (len(y), 11)
    assert np.isfinite(x).all()
    assert set(np.unique(y)) <= {0.0, 1.0}
    # both classes are present: candidates next to the beam (within 4 cells
    # of a cell above the threshold) fill up, those further out stay empty
    assert 0.5 < y.mean() < 1.0 and 0.0 < near.mean() < 1.0
    assert y[~near].mean() > y[near].mean()
    # own max at or below the tile threshold (feature: log10(max phi) / 10)
    lg_thr = np.log10(cfg.device.tile_threshold) / 10
    assert (x[:, 0] >= -3.0).all() and (x[:, 0] <= lg_thr + 1e-6).all()
    assert (x[:, 2] >= 0).all() and (x[:, 2] <= 1).all()
    assert (x[:, 3] > 0.2).all() and (x[:, 3] < 2.0).all()       # g/cm^3
    assert (x[:, 4] > 0).all() and (x[:, 4] <= 1).all()


@pytest.fixture(scope="module")
def trained(samples):
    """A network trained on 80 % of the samples; the rest are held out (about
    12,000 candidates, so that one false skip is below 1e-4 of them)."""
    x, y, near = samples
    order = np.random.default_rng(0).permutation(len(y))
    n_test = len(y) // 5
    fit, test = order[n_test:], order[:n_test]
    model, loss = TS.train(x[fit], y[fit], epochs=20, batch=2048, lr=1e-3)
    say("training loss", loss)
    return model, (x[test], y[test], near[test])


def test_network_beats_the_fixed_guard_rule(trained):
    """On held-out candidates: the fixed rule leaves out those with nothing
    above the threshold within 4 cells (no false skips by construction);
    the network, at a decision probability with no more false skips than
    1e-4 of the candidates, leaves out at least 1.5 times as many."""
    model, (x, y, near) = trained
    report = TS.against_guard_rule(model, x, y, near,
                                   decisions=(0.9, 0.99, 0.995, 0.999, 0.9999, 0.99999))
    for r in report["rules"]:
        say(r)
    rule = next(r for r in report["rules"] if r["rule"] == "guard 4 cells")
    nets = [r for r in report["rules"] if r["rule"].startswith("network")]
    good = [r for r in nets if r["false_skip_fraction"] <= 1e-4]
    assert good, "no decision level meets the false-skip bound"
    best = max(good, key=lambda r: r["skip_fraction"])
    assert best["skip_fraction"] >= 1.5 * rule["skip_fraction"]


def test_trained_network_skips_tiles_and_holds_the_dose(trained, target):
    model, _ = trained
    handle, num, dose0, steps0, tiles0, per0 = target
    dose, steps, tiles, judged, rejections, back, _ = _march(
        handle, num, model, TS.SkipOptions(decision=0.9, infer_every=1))
    change = np.abs(dose - dose0).max() / dose0.max()
    say("network: tiles", tiles, "vs", tiles0, "ratio", tiles / tiles0, "judged", judged,
        "rejections", rejections, "dose change / max", change)
    # the fixed 4-cell rule on the same target: every candidate it admits left out
    d4, s4, tiles4, *_ = _march(handle, num, _constant_model(30.0),
                                TS.SkipOptions(guard=4, judge_every=0))
    say("fixed 4-cell guard rule: tiles", tiles4, "ratio", tiles4 / tiles0)
    assert judged > 0
    assert tiles < 0.9 * tiles0 and tiles < tiles4
    assert change < 1e-3


def test_never_skipping_network_is_bitwise_the_plain_march(target):
    handle, num, dose0, steps0, tiles0, per0 = target
    dose, steps, tiles, judged, rejections, back, _ = _march(
        handle, num, _constant_model(-30.0), TS.SkipOptions(judge_every=4))
    say("never-skip network: judged", judged, "tiles", tiles, "vs", tiles0)
    assert judged > 0 and rejections == 0 and (back == -1).all()
    assert np.array_equal(steps, steps0)
    assert np.array_equal(dose, dose0)


def test_graph_steps_with_the_network_are_bitwise_the_launched_steps(target):
    """Network on: plain steps run as graphs (the network gated on the
    device), judged steps as separate launches; same march bit for bit."""
    handle, num, dose0, steps0, tiles0, per0 = target
    model = TS.load_model(0) if TS.WEIGHTS_FILE.exists() else _constant_model(30.0)
    opts = TS.SkipOptions(judge_every=16)
    d_launch, s_launch, *_ = _march(handle, num, model, opts)
    handle.set_graphs(True)
    try:
        d_graph, s_graph, tiles, judged, rejections, back, _ = _march(handle, num, model, opts)
        n_graph, err = handle.graph_info()
    finally:
        handle.set_graphs(False)
    say("network with graphs: graph steps", n_graph, "judged", judged, "error", repr(err))
    assert err == "" and n_graph > 0 and judged > 0
    assert np.array_equal(s_launch, s_graph)
    assert np.array_equal(d_launch, d_graph)


def test_judge_rolls_back_an_always_skipping_network(target):
    handle, num, dose0, steps0, tiles0, per0 = target
    dose, steps, tiles, judged, rejections, back, per = _march(
        handle, num, _constant_model(30.0),
        TS.SkipOptions(judge_every=8, judge_tol=1e-12, guard=4))
    say("always-skip network, tol 1e-12: judged", judged, "rejections", rejections,
        "rolled back to", back.tolist())
    assert rejections > 0 and (back >= 0).sum() == rejections
    for b, (got, ref) in enumerate(zip(per, per0)):
        if back[b] == 0:
            assert np.array_equal(got, ref), b
            assert steps[b] == steps0[b]
        assert np.abs(got - ref).max() <= 1e-3 * ref.max(), b
    # loose tolerance: the same network is accepted and the dose still holds
    dose, steps, tiles, judged, rejections, back, _ = _march(
        handle, num, _constant_model(30.0), TS.SkipOptions(judge_every=8, guard=4))
    change = np.abs(dose - dose0).max() / dose0.max()
    say("always-skip network, tol 1e-3: rejections", rejections, "tiles", tiles / tiles0,
        "dose change / max", change)
    assert change < 1e-3


def test_argument_checks(target):
    handle = target[0]
    model = _constant_model(0.0)
    for kw in ({"decision": 0.0}, {"decision": 1.0}, {"horizon": 0}, {"horizon": 256},
               {"infer_every": 0}, {"infer_every": 9}, {"judge_every": -1},
               {"judge_tol": 0.0}):
        with pytest.raises(ValueError):
            handle.set_skip_model(model, **kw)
    with pytest.raises(ValueError):
        handle.record(True, horizon=0)
    with pytest.raises(ValueError):
        handle.record(True, capacity=0)
    handle.set_skip_model(None)
    x, y = handle.samples()
    assert x.shape == (0, 11) and y.shape == (0,)
