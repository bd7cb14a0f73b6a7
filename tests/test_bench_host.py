"""bench.py's host-side accounting, CPU only.

  * a request for more than one GPU exits with status 2 and a message;
  * the algorithmic bytes per cell of each stage kernel follow from the
    arrays the stage reads and writes (state of 4 moments, material id of
    1 byte, dose read-modify-write);
  * the roofline entry is the stage with the largest total time, its
    achieved bandwidth bytes x cells / average launch time;
  * the wall-clock budget cuts the warm-up and timed batches so that the
    driver's `--steps 20 --warmup 5` finishes in time;
  * the CPU baseline (the oracle on a few energy steps of a box) reports the
    all-thread and the one-thread rate with the host CPU model;
  * the combined workload's FP64 samples are air-free beamlets that stop
    inside the phantom.
"""

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""


def _bytes(word, state_reads, scalar_reads, state_writes, dose_rmw):
    """Bytes per cell: 4-moment state arrays read and written, scalar reads
    (V1.phi), the 1-byte material id and the dose read-modify-write."""
    return 4 * word * (state_reads + state_writes) + word * scalar_reads + 1 + dose_rmw


def test_stage_bytes_follow_from_the_arrays_moved():
    # stage 0: U -> V1; stage 1: V1, U -> V2; stage 2: V2, U, V1.phi -> U and
    # the dose (FP64: one double read and written; FP32 bulk: a compensated
    # double pair, value and compensation read and written)
    fp64 = {"stage0": _bytes(8, 1, 0, 1, 0), "stage1": _bytes(8, 2, 0, 1, 0),
            "stage2": _bytes(8, 2, 1, 1, 2 * 8)}
    assert fp64 == {"stage0": 65, "stage1": 97, "stage2": 121} == bench.STAGE_BYTES_FP64
    fp32 = {"stage0": _bytes(4, 1, 0, 1, 0), "stage1": _bytes(4, 2, 0, 1, 0),
            "stage2": _bytes(4, 2, 1, 1, 2 * 2 * 8)}
    assert fp32 == bench.STAGE_BYTES_FP32 == bench.STAGE_BYTES_FP32_SINGLE
    # per cell-step: B = 33 s + 3 + 16 (1 + K), K = 0 in FP64, 1 with the pair
    assert bench.STEP_BYTES_FP64 == 33 * 8 + 3 + 16 == 283
    assert bench.STEP_BYTES_FP32 == 33 * 4 + 3 + 32 == 167


def test_whole_step_roofline():
    r = bench.whole_step_roofline(2e10, 4.0, 167)
    assert r["achieved"] == pytest.approx(2e10 * 167 / 4.0 / 1e9, abs=0.01)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, abs=1e-5)
    assert bench.whole_step_roofline(1.0, 0.0, 167) is None


def test_budget_fits_the_drivers_command():
    """`--steps 20 --warmup 5` with a 17.6 s batch (the 256-beamlet batch's
    time) is cut to fit: one profiled warm-up, a second one only when the
    timed region keeps its length, timed batches within the timed cap and the
    whole budget; what remains covers the untimed legs."""
    batch_s, budget, cap = 17.6, 450.0, 60.0
    reserve = 1.2 * batch_s + 60.0 + 30.0
    spent = 45.0 + batch_s           # setup and autotune, then the profiled warm-up
    warm = 1
    while bench.plan_extra_warmup(5, warm, batch_s, spent, budget, reserve, cap):
        spent += batch_s
        warm += 1
    assert warm == 2
    k = bench.plan_timed_steps(20, batch_s, spent, budget, reserve, cap)
    assert k == 3 and k * batch_s <= cap
    assert spent + k * batch_s + reserve <= budget
    # a slow batch still times one; a fast one stops at the request
    assert bench.plan_timed_steps(20, 500.0, spent, budget, reserve, cap) == 1
    assert bench.plan_timed_steps(4, 0.5, spent, budget, reserve, cap) == 4
    # little budget left: no extra warm-up
    assert not bench.plan_extra_warmup(5, 1, batch_s, 400.0, budget, reserve, cap)
    assert not bench.plan_extra_warmup(1, 1, 1.0, 0.0, budget, 0.0, cap)


def test_fp64_samples_and_batch_rate():
    energies = [70.0 + i for i in range(11)]
    picks = bench.fp64_sample_order(energies, 3)
    assert picks == [5, 0, 10]
    assert len(set(bench.fp64_sample_order(energies, 6))) == 6
    rows = [{"fp64_seconds": 10.0, "fp64_energy_steps": 400, "batch_energy_steps": 400},
            {"fp64_seconds": 50.0, "fp64_energy_steps": 2000, "batch_energy_steps": 1800}]
    steps = np.array([400, 1000, 1800])
    rate, sps, ratio = bench.fp64_batch_rate(rows, steps)
    assert sps == pytest.approx(0.025)
    assert ratio == pytest.approx(0.5 * (1.0 + 2000 / 1800))
    assert rate == pytest.approx(3 / (0.025 * ratio * 3200))


def test_roofline_takes_the_longest_stage():
    stats = {"stage0": (10, 12.0), "stage1": (10, 15.0), "stage2": (8, 16.0), "cfl": (10, 99.0)}
    cells = 1_000_000
    r = bench.roofline_of(stats, bench.STAGE_BYTES_FP32, lambda launches: cells,
                          {"stage2": 136.8}, {"stage2": 0.75})
    assert r["kernel"] == "stage2" and r["avg_launch_ms"] == 2.0
    achieved = 85 * cells / 2e-3 / 1e9
    assert r["achieved"] == pytest.approx(achieved, abs=0.01)
    assert r["frac"] == pytest.approx(achieved / 8000.0, abs=1e-5)
    assert r["traffic"] == round(136.8 * cells) and r["valu_issue_frac_profiled"] == 0.75
    assert r["traffic_profile"].startswith("profiles/")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert bench.roofline_of({}, bench.STAGE_BYTES_FP32, lambda n: cells) is None


def test_cpu_baseline_of_a_batch_beamlet():
    from proton_solver.batch import batch_workload, beamlet_boxes

    phantom, beamlets = batch_workload(n_beamlets=4)
    box = beamlet_boxes(phantom, beamlets[:1])[0]
    r = bench.batch_cpu_baseline(phantom, beamlets[0], box, bench.batch_config(), 2, 1e9)
    assert r["kind"] == "port" and r["unit"] == "beamlet dose solves/s" and r["cores"] >= 1
    assert r["value"] == pytest.approx(r["cell_steps_per_s"] / 1e9)
    assert "first 2 energy steps" in r["sample"]
    assert r["single_thread"]["cell_steps_per_s"] > 0
    assert "first 1 energy steps" in r["single_thread"]["sample"]
    assert r["cpu_model"] is None or isinstance(r["cpu_model"], str)


def test_combined_samples_are_air_free_beamlets_that_stop_inside():
    """The combined workload's FP64 samples: air-free beamlets whose boxes end
    before the far face (an exiting beam's moving domain reaches air past
    it), spread over the energies, among them one above 150 MeV."""
    from proton_solver.batch import (air_free_beamlets, batch_workload_config, beamlet_boxes,
                                     combined_workload)

    phantom, beamlets = combined_workload()
    n = batch_workload_config().numerics
    boxes = beamlet_boxes(phantom, beamlets, n.energy_min, n.energy_window_sigma)
    air_free = air_free_beamlets(phantom, beamlets, boxes)
    picks = bench.combined_sample_picks(phantom, beamlets, boxes, air_free, 3)
    assert len(picks) == 3
    assert set(picks) <= set(air_free)
    assert all(boxes[b][4] < phantom.shape[2] for b in picks)
    # some air-free beamlets exit through the far face and are left out
    assert any(boxes[b][4] == phantom.shape[2] for b in air_free)
    energies = sorted(beamlets[b].energy for b in picks)
    assert energies[0] < 80.0 and energies[-1] > 150.0
    assert bench.combined_sample_picks(phantom, beamlets, boxes, air_free, 0) == []
