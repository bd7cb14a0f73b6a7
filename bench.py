"""Benchmark of the MI355X M1 energy march (one JSON line on stdout).

A step is one beamlet dose solve: a full energy march of the workload below,
from E0 = E_beam + 5 sigma_E down to energy_min, with the problem (material
ids, tables, inflow profile) resident in HBM before the timed region.

Workload (the 150 MeV single-beamlet configuration of the metric): water,
0.5 mm cells, 120 x 120 x 360 grid (6 x 6 x 18 cm), Gaussian beam sigma 4 mm,
sigma_E 1 MeV, Wentzel-VI scattering + nuclear removal, IF-SSPRK3, WENO-Z,
exact M1 speeds, cfl 0.9, max energy step 0.25 MeV, energy_min 1 MeV.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32_bulk|fp64|fp64_fast]

The measured march is FP32-bulk by default (float reconstruction, flux and
stage algebra; double closure, projection and dose accumulation); one plain
FP64 solve of the same beamlet, outside the timed region, states its
max-dose error and gamma(1%/1 mm) pass rate and the speed-up over FP64.

Reported: beamlet dose solves/s (value), cell-step updates/s, the HBM
roofline of the dominant stage kernel (algorithmic bytes per launch over its
HIP-event duration on the march stream) and the CPU oracle timed on a bounded
sample of the same workload on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "proton-dose-m1_amd"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
MEV = 1.0

# Algorithmic HBM bytes per cell of each FP64 stage kernel: every stage input
# read once, every output written once, material id 1 byte:
#   stage 0: read U (32) + ids (1), write V1 (32)
#   stage 1: read V1 (32) + U (32) + ids (1), write V2 (32)
#   stage 2: read V2 (32) + U (32) + V1.phi (8) + ids (1), write U (32), dose RMW (16)
STAGE_BYTES_FP64 = {"stage0": 65, "stage1": 97, "stage2": 121}
# FP32-bulk: float state (16 B per cell-state), dose as a compensated float pair:
#   stage 0: read U (16) + ids (1), write V1 (16)
#   stage 1: read V1 (16) + U (16) + ids (1), write V2 (16)
#   stage 2: read V2 (16) + U (16) + V1.phi (4) + ids (1), write U (16), dose pair RMW (16)
STAGE_BYTES_FP32 = {"stage0": 33, "stage1": 49, "stage2": 69}


def workload_config(precision):
    from proton_solver import (
        Beam, DeviceOptions, EnergyScheme, Geometry, Numerics, Physics, Scattering,
        SolverConfig, SpatialScheme, WaveSpeeds,
    )

    return SolverConfig(
        geometry=Geometry(size=(0.06, 0.06, 0.18), resolution=(120, 120, 360)),
        beam=Beam(n_protons=1e6, energy=150.0, energy_spread=1.0, sigma=(0.004, 0.004),
                  center=(0.0, 0.0), angular_spread=0.0, cutoff_sigma=5.0),
        physics=Physics(scattering=Scattering.WENTZEL, nuclear=True),
        numerics=Numerics(spatial_scheme=SpatialScheme.WENO_Z,
                          energy_scheme=EnergyScheme.IF_SSPRK3, wave_speeds=WaveSpeeds.EXACT,
                          max_energy_step=0.25, cfl=0.9, energy_min=1.0),
        device=DeviceOptions(precision=precision),
    )


def host_problem(cfg):
    from proton_solver.numerics.boundary import BeamInflow
    from proton_solver.numerics.grid import Grid
    from proton_solver.physics.materials import MaterialMap

    grid = Grid.from_geometry(cfg.geometry, np)
    mm = MaterialMap.from_geometry(cfg.geometry, grid, np)
    inflow = BeamInflow.from_beam(cfg.beam, grid, np)
    return grid, mm, inflow


def cpu_baseline(cfg, n_steps_full, sample_steps):
    """Oracle (C/OpenMP FP64 restatement) on the first energy steps of the march."""
    from oracle.oracle import OracleProblem

    grid, mm, inflow = host_problem(cfg)
    op = OracleProblem(grid, mm, inflow, cfg)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    op.march(max_steps=1)  # warm-up (page-in, thread pool)
    t0 = time.perf_counter()
    _, _, E, _ = op.march(max_steps=sample_steps)
    dt = time.perf_counter() - t0
    steps = len(E) - 1
    n = int(np.prod(grid.shape))
    per_step = dt / steps
    return {
        "value": 1.0 / (per_step * n_steps_full),
        "unit": "beamlet dose solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {steps} of {n_steps_full} energy steps of the same "
                   f"{grid.shape[0]}x{grid.shape[1]}x{grid.shape[2]} march "
                   f"({dt:.1f} s), extrapolated to the full march"),
        "cell_steps_per_s": n * steps / dt,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32_bulk", choices=["fp32_bulk", "fp64", "fp64_fast"])
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="skip the FP64 solve used to state the error of the measured march")
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: the march of one beamlet runs on one GPU and "
              "this build has no multi-GPU path (beamlet sharding is described in DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    from proton_solver import Precision
    from proton_solver import _native

    precision = {"fp64": Precision.FP64, "fp64_fast": Precision.FP64_FAST,
                 "fp32_bulk": Precision.FP32_BULK}[args.precision]
    cfg = workload_config(precision)
    grid, mm, inflow = host_problem(cfg)
    n_cells = int(np.prod(grid.shape))

    ctx = _native.Context(0)
    ctx.set_problem(grid, mm, inflow, cfg.beam)
    num = _native.numerics_struct(cfg)
    stage_bytes = STAGE_BYTES_FP32 if precision is Precision.FP32_BULK else STAGE_BYTES_FP64

    # Plain FP64 march of the same beamlet: the accuracy reference of the
    # measured march and the denominator of its speed-up (outside the timed region).
    fp64 = None
    if not args.no_fp64_reference and precision is not Precision.FP64:
        num64 = _native.numerics_struct(cfg, _native.FP64)
        ctx.set_state(None)
        t0 = time.perf_counter()
        E64, _ = ctx.march(num64)
        fp64 = {"seconds": time.perf_counter() - t0, "steps": len(E64) - 1,
                "dose": ctx.get_dose()}

    for _ in range(args.warmup):
        ctx.set_state(None)
        ctx.march(num)
    ctx.set_profiling(True)
    ctx.reset_stats()
    steps_per_march = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.set_state(None)
        E, _ = ctx.march(num)
        steps_per_march.append(len(E) - 1)
    elapsed = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    tiles_eval, tiles_full = ctx.tile_stats()
    ctx.set_profiling(False)
    dose = ctx.get_dose()

    n_energy_steps = steps_per_march[-1]
    ms_per_step = 1000.0 * elapsed / args.steps
    solves_per_s = args.steps / elapsed
    cell_steps = n_cells * sum(steps_per_march)

    # Dominant stage kernel and its HBM roofline.  Cells per launch are the
    # cells of the evaluated tiles (all cells when no tile is skipped).
    eval_frac = (tiles_eval / tiles_full) if tiles_full else 1.0
    cells_per_launch = n_cells * eval_frac
    best = None
    for name, per_cell in stage_bytes.items():
        launches, total_ms = stats.get(name, (0, 0.0))
        if launches == 0:
            continue
        avg_s = total_ms / launches / 1000.0
        achieved = per_cell * cells_per_launch / avg_s / 1e9
        if best is None or total_ms > best[1]:
            best = (name, total_ms, achieved, avg_s, per_cell)
    roofline = None
    if best is not None:
        name, total_ms, achieved, avg_s, per_cell = best
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None,
            "kernel": name, "avg_launch_ms": round(avg_s * 1000, 4),
            "algorithmic_bytes_per_cell": per_cell,
            "cells_per_launch": round(cells_per_launch),
        }

    accuracy = None
    if fp64 is not None:
        from proton_solver import metrics

        ref = fp64["dose"]
        accuracy = {
            "reference": "plain FP64 march on the same GPU (bit-identical to the CPU oracle)",
            "max_dose_rel_err": metrics.maximum_dose_relative_difference(dose, ref),
            "max_abs_diff_over_max": metrics.max_abs_difference_over_max(dose, ref),
            "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(
                dose, ref, (grid.dx, grid.dy, grid.dz), 1.0, 1.0, 10.0),
            "fp64_solve_s": fp64["seconds"],
            "fp64_energy_steps": fp64["steps"],
            "speedup_vs_fp64_march": fp64["seconds"] / (elapsed / args.steps),
        }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, n_energy_steps, args.cpu_sample_steps)

    line = {
        "metric": "beamlet dose solves/s at <=0.5% max-dose err vs plain FP64; cell-step updates/s",
        "value": solves_per_s,
        "unit": "beamlet dose solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64_fast": "f64",
                  "fp32_bulk": "f32 bulk / f64 closure+projection+dose"}[args.precision],
        "data": "synthetic water phantom (no external data)",
        "config": {
            "workload": "150 MeV water pencil beam, 0.5 mm, 120x120x360, Wentzel+nuclear, "
                        "IF-SSPRK3, one beamlet per step",
            "precision": args.precision,
            "grid": list(grid.shape),
            "energy_steps_per_solve": n_energy_steps,
            "max_dose_Gy": float(dose.max()),
        },
        "cell_steps_per_s": cell_steps / elapsed,
        "cell_steps_per_s_evaluated": cell_steps * eval_frac / elapsed,
        "tiles_evaluated_fraction": eval_frac,
        "kernel_ms_per_solve": {k: round(v[1] / max(args.steps, 1), 3) for k, v in stats.items()
                                if v[0]},
        "accuracy_vs_fp64": accuracy,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
