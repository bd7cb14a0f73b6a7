"""bench.py's host-side accounting, CPU only.

  * a request for more than one GPU exits with status 2 and a message;
  * the algorithmic bytes per cell of each stage kernel follow from the
    arrays the stage reads and writes (state of 4 moments, material id of
    1 byte, dose read-modify-write);
  * the roofline entry is the stage with the largest total time, its
    achieved bandwidth bytes x cells / average launch time;
  * the CPU baseline (the oracle on a few energy steps of a box) reports the
    all-thread and the one-thread rate with the host CPU model.
"""

import subprocess
import sys
from pathlib import Path

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
    # the dose (FP64: one double read and written; float: a compensated pair)
    fp64 = {"stage0": _bytes(8, 1, 0, 1, 0), "stage1": _bytes(8, 2, 0, 1, 0),
            "stage2": _bytes(8, 2, 1, 1, 2 * 8)}
    assert fp64 == {"stage0": 65, "stage1": 97, "stage2": 121} == bench.STAGE_BYTES_FP64
    single = {"stage0": _bytes(4, 1, 0, 1, 0), "stage1": _bytes(4, 2, 0, 1, 0),
              "stage2": _bytes(4, 2, 1, 1, 2 * 2 * 4)}
    assert single == bench.STAGE_BYTES_FP32_SINGLE
    # the batch: stage 0 also writes W1, stage 1 reads W1 and writes W2,
    # stage 2 reads W2 (the stencil inputs formed once per cell)
    batch = {"stage0": _bytes(4, 1, 0, 2, 0), "stage1": _bytes(4, 3, 0, 2, 0),
             "stage2": _bytes(4, 3, 1, 1, 2 * 2 * 4)}
    assert batch == bench.STAGE_BYTES_FP32


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
