"""Benchmark of the MI355X M1 energy march (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch256|single]

Workload "batch256" (default): 256 pencil beamlets marched as one batch
through a 200 x 200 x 300 water phantom at 1 mm holding 12 seeded ellipsoids
(semi-axes 5-30 mm) of cortical bone (1.85 g/cm^3) or lung (0.26 g/cm^3);
beamlet energies U(70, 230) MeV, sigma U(2, 6) mm, entry points U(-30, 30) mm,
each beamlet on its own 64 x 64 x depth-to-range box; Wentzel-VI scattering +
nuclear removal, IF-SSPRK3, WENO-Z, exact M1 speeds, cfl 0.9, max energy
step 0.25 MeV, energy_min 1 MeV, FP32-bulk arithmetic with tile skipping.
A step is one batch (256 beamlet dose solves); value is beamlets/s.  Outside
the timed region a sample of the beamlets is marched by the plain FP64
full-domain march (the whole phantom; bitwise the CPU oracle): its time per
beamlet gives the FP64 beamlets/s and the speed-up, and its doses the
max-dose error and gamma(1 %/1 mm) pass rate of the batch.

Workload "single": one 150 MeV beamlet in water, 0.5 mm cells,
120 x 120 x 360 grid, sigma 4 mm, same physics and numerics, FP32-bulk
(--precision selects fp64 / fp64_fast); a step is one beamlet solve.

Reported with either: cell-step updates/s, the HBM roofline of the dominant
stage kernel (algorithmic bytes per launch over its HIP-event duration on the
march stream) and the CPU oracle timed on a bounded sample of the same
workload on this host.
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
# FP32-bulk batch: float state (16 B per cell-state), dose as a compensated
# float pair, the next stage's projected stencil input W formed by stages 0, 1:
#   stage 0: read U (16) + ids (1), write V1 (16) + W1 (16)
#   stage 1: read W1 (16) + V1 (16) + U (16) + ids (1), write V2 (16) + W2 (16)
#   stage 2: read W2 (16) + V2 (16) + U (16) + V1.phi (4) + ids (1), write U (16),
#            dose pair RMW (16)
STAGE_BYTES_FP32 = {"stage0": 49, "stage1": 81, "stage2": 85}
# FP32-bulk single box (stages 1, 2 form their input at load, no W):
#   stage 0: read U + ids, write V1; stage 1: read V1 + U + ids, write V2;
#   stage 2: read V2 + U + V1.phi + ids, write U, dose pair RMW
STAGE_BYTES_FP32_SINGLE = {"stage0": 33, "stage1": 49, "stage2": 69}
# HBM bytes per evaluated cell of the batched stage kernels per tile shape,
# from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the same counters
# on a kernel of known traffic (profiles/stage_kernels_batch_pmc_traffic.md);
# shapes without a measurement report traffic null.
PMC_BYTES_PER_CELL_FP32 = {
    (8, 16, 16): {"stage0": 87.6, "stage1": 132.3, "stage2": 136.8},
}
# Fraction of the VALU issue slots the batched stage kernels use, per tile
# shape (rocprofv3 SQ_INSTS_VALU x 4 cycles / (kernel time x 2.4 GHz x 1024
# SIMDs), profiles/stage_kernels_batch_valu.md): the bound of these kernels.
VALU_ISSUE_FRAC_FP32 = {
    (8, 16, 16): {"stage0": 0.86, "stage1": 0.77, "stage2": 0.75},
}


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


def host_cpu_model():
    """Model name of the host CPU (from /proc/cpuinfo), or None."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def single_thread_rate(op, n_cells, steps, threads):
    """Cell-steps/s of the same oracle march on one host thread (a shorter
    sample), then the thread count restored."""
    from oracle.oracle import set_threads

    set_threads(1)
    try:
        t0 = time.perf_counter()
        _, _, E, _ = op.march(max_steps=steps)
        dt = time.perf_counter() - t0
    finally:
        set_threads(threads)
    done = len(E) - 1
    return {"cell_steps_per_s": n_cells * done / dt,
            "sample": f"first {done} energy steps ({dt:.1f} s)"}


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
    one = single_thread_rate(op, n, max(sample_steps // 10, 1), threads)
    return {
        "value": 1.0 / (per_step * n_steps_full),
        "unit": "beamlet dose solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {steps} of {n_steps_full} energy steps of the same "
                   f"{grid.shape[0]}x{grid.shape[1]}x{grid.shape[2]} march "
                   f"({dt:.1f} s), extrapolated to the full march"),
        "cell_steps_per_s": n * steps / dt,
        "cpu_model": host_cpu_model(),
        "single_thread": one,
    }


def roofline_of(stats, stage_bytes, cells_per_launch, traffic_per_cell=None, valu_frac=None):
    """HBM roofline of the stage kernel with the largest total time; traffic
    is the measured HBM bytes per launch when per-cell figures are given, and
    valu_issue_frac_profiled the kernel's profiled VALU issue fraction."""
    best = None
    for name, per_cell in stage_bytes.items():
        launches, total_ms = stats.get(name, (0, 0.0))
        if launches == 0:
            continue
        avg_s = total_ms / launches / 1000.0
        achieved = per_cell * cells_per_launch(launches) / avg_s / 1e9
        if best is None or total_ms > best[1]:
            best = (name, total_ms, achieved, avg_s, per_cell, cells_per_launch(launches))
    if best is None:
        return None
    name, total_ms, achieved, avg_s, per_cell, cells = best
    return {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 5),
        "traffic": (round(traffic_per_cell[name] * cells) if traffic_per_cell else None),
        "kernel": name, "avg_launch_ms": round(avg_s * 1000, 4),
        "algorithmic_bytes_per_cell": per_cell, "cells_per_launch": round(cells),
        "valu_issue_frac_profiled": (valu_frac or {}).get(name),
    }


def batch_config(tile_threshold=None):
    from proton_solver.batch import BATCH_TILE_THRESHOLD, batch_workload_config

    return batch_workload_config(BATCH_TILE_THRESHOLD if tile_threshold is None
                                 else tile_threshold)


def batch_cpu_baseline(phantom, beamlet, box, cfg, sample_steps, cell_steps_per_beamlet):
    """Oracle (C/OpenMP FP64) on the first steps of one beamlet's box march."""
    from oracle.oracle import OracleProblem
    from proton_solver.batch import box_problem

    grid, mm, inflow, beam = box_problem(phantom, box, beamlet)
    cfg1 = type(cfg)(geometry=cfg.geometry, beam=beam, physics=cfg.physics,
                     numerics=cfg.numerics, device=cfg.device)
    op = OracleProblem(grid, mm, inflow, cfg1)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    op.march(max_steps=1)
    t0 = time.perf_counter()
    _, _, E, _ = op.march(max_steps=sample_steps)
    dt = time.perf_counter() - t0
    steps = len(E) - 1
    n = int(np.prod(grid.shape))
    rate = n * steps / dt
    one = single_thread_rate(op, n, max(sample_steps // 10, 1), threads)
    return {
        "value": rate / cell_steps_per_beamlet,
        "unit": "beamlet dose solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {steps} energy steps of one {beamlet.energy:.1f} MeV beamlet on its "
                   f"{grid.shape[0]}x{grid.shape[1]}x{grid.shape[2]} box ({dt:.1f} s, "
                   f"{rate:.3g} cell-steps/s), extrapolated with the batch's mean box "
                   f"cell-steps per beamlet"),
        "cell_steps_per_s": rate,
        "cpu_model": host_cpu_model(),
        "single_thread": one,
    }


def run_batch(args):
    from proton_solver import _native, metrics
    from proton_solver.batch import (
        batch_workload, beamlet_boxes, full_problem, prepare_batch,
    )

    cfg = batch_config(args.tile_threshold)
    phantom, beamlets = batch_workload(n_beamlets=args.beamlets)
    # Tile shape and launch from the autotuner (a sample of the beamlets,
    # outside the timed region).
    tuned = None
    if args.autotune and args.tile_shape is None:
        from proton_solver.autotune import autotune

        t0 = time.perf_counter()
        choice = autotune(phantom, beamlets, cfg, sample=args.autotune_sample,
                          modes=("per_tile", "persistent"),
                          log=lambda row: print(json.dumps({"autotune": row}), file=sys.stderr,
                                                flush=True))
        tuned = {"shape": list(choice.shape), "mode": choice.mode,
                 "sample_beamlets": args.autotune_sample,
                 "seconds": round(time.perf_counter() - t0, 2), "table": choice.table}
    t0 = time.perf_counter()
    boxes = beamlet_boxes(phantom, beamlets, cfg.numerics.energy_min,
                          cfg.numerics.energy_window_sigma)
    handle, boxes = prepare_batch(phantom, beamlets, cfg, boxes,
                                  tile_shape=(choice.shape_index if tuned else
                                              (args.tile_shape or 0)))
    if tuned:
        choice.apply(handle)
    setup_s = time.perf_counter() - t0
    num = _native.numerics_struct(cfg)
    from proton_solver import tileskip

    use_net = args.skip_network == "on" or (args.skip_network == "auto" and
                                            tileskip.WEIGHTS_FILE.exists())
    net_opts = tileskip.SkipOptions()
    model = tileskip.load_model(0) if use_net else None
    tileskip.enable(handle, model, net_opts)
    for _ in range(args.warmup):
        handle.march(num, want_dose=False)
    handle.set_profiling(True)
    handle.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dose, steps = handle.march(num, want_dose=False)
    elapsed = time.perf_counter() - t0
    stats = handle.kernel_stats()
    ev, full = handle.tile_stats()
    host_ms = handle.host_ms()
    handle.set_profiling(False)
    K = max(args.steps, 1)
    B = len(beamlets)

    # The network's effect: the same batch marched with the dilation rule alone
    # (outside the timed region).
    net = None
    beamlet_dose = handle.beamlet_dose
    if use_net:
        judged, rejections, back = handle.skip_stats()
        net_doses = [handle.beamlet_dose(b) for b in range(B)]
        beamlet_dose = net_doses.__getitem__
        net = {"weights": tileskip.WEIGHTS_FILE.name, "decision": net_opts.decision,
               "horizon": net_opts.horizon, "infer_every": net_opts.infer_every,
               "judge_every": net_opts.judge_every,
               "judge_tol": net_opts.judge_tol, "judged_steps": judged,
               "rejected_beamlets": rejections}
        if not args.no_network_reference:
            tileskip.enable(handle, None)
            handle.reset_stats()
            t1 = time.perf_counter()
            _, steps0 = handle.march(num, want_dose=False)
            sec0 = time.perf_counter() - t1
            ev0, _ = handle.tile_stats()
            worst = 0.0
            for b in range(B):
                ref = handle.beamlet_dose(b)
                worst = max(worst, float(np.abs(net_doses[b] - ref).max() / ref.max()))
            net.update({
                "dilation_only_seconds": round(sec0, 3),
                "dilation_only_beamlets_per_s": B / sec0,
                "tiles_vs_dilation_only": ev / K / ev0,
                "worst_beamlet_max_dose_change_over_max": worst,
            })
        print(json.dumps({"tile_skip_network": net}), file=sys.stderr, flush=True)
    box_cells = np.array([b[2] * b[3] * b[4] for b in boxes], dtype=np.float64)
    tx, ty, tz = handle.tile_shape
    tile_cap = np.array([-(-b[2] // tx) * -(-b[3] // ty) * -(-b[4] // tz) * (tx * ty * tz)
                         for b in boxes], dtype=np.float64)
    fill = box_cells.sum() / tile_cap.sum()
    cell_steps = float((box_cells * steps).sum()) * K
    n_phantom = int(np.prod(phantom.shape))
    eval_frac = ev / full if full else 1.0

    def cells_per_launch(launches):
        return ev * float(tx * ty * tz) * fill / launches

    roof = roofline_of(stats, STAGE_BYTES_FP32, cells_per_launch,
                       PMC_BYTES_PER_CELL_FP32.get(tuple(handle.tile_shape)),
                       VALU_ISSUE_FRAC_FP32.get(tuple(handle.tile_shape)))
    if args.profile_calibration:
        # a kernel of known traffic for scaling the HBM counters: the beamlet
        # dose read-out reads a float pair and writes a double per box cell
        x0, y0, nx, ny, nz = boxes[0]
        handle.beamlet_dose(0)
        print(json.dumps({"calibration_kernel": "box_dose_kernel", "cells": nx * ny * nz,
                          "bytes_read": 8 * nx * ny * nz, "bytes_written": 8 * nx * ny * nz}),
              file=sys.stderr, flush=True)

    # Plain FP64 full-domain march of a sample of the beamlets (outside the timed region).
    fp64 = None
    if args.fp64_sample > 0:
        order = np.argsort([b.energy for b in beamlets])
        picks = [int(order[int(round(q * (B - 1)))])
                 for q in np.linspace(0.0, 1.0, args.fp64_sample)]
        ctx = _native.Context(0)
        num64 = _native.numerics_struct(cfg, _native.FP64)
        rows = []
        for b in picks:
            grid, mm, inflow, beam = full_problem(phantom, beamlets[b])
            ctx.set_problem(grid, mm, inflow, beam)
            t1 = time.perf_counter()
            E64, _ = ctx.march(num64)
            sec = time.perf_counter() - t1
            ref = ctx.get_dose()
            x0, y0, nx, ny, nz = boxes[b]
            got = np.zeros(phantom.shape)
            got[x0:x0 + nx, y0:y0 + ny, :nz] = beamlet_dose(b)
            # gamma on the box plus a 4-cell margin (the FP64 dose beyond it is < 1e-3 of max)
            sl = (slice(max(x0 - 4, 0), x0 + nx + 4), slice(max(y0 - 4, 0), y0 + ny + 4),
                  slice(0, min(nz + 4, phantom.shape[2])))
            outside = ref.copy()
            outside[sl] = 0.0
            rows.append({
                "beamlet": b, "energy_MeV": round(beamlets[b].energy, 3),
                "fp64_seconds": round(sec, 3), "fp64_energy_steps": len(E64) - 1,
                "batch_energy_steps": int(steps[b]),
                "max_dose_rel_err": metrics.maximum_dose_relative_difference(got, ref),
                "max_abs_diff_over_max": metrics.max_abs_difference_over_max(got, ref),
                "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(
                    got[sl], ref[sl], phantom.spacing, 1.0, 1.0, 10.0),
                "fp64_dose_outside_gamma_region_over_max": float(outside.max() / ref.max()),
            })
            print(json.dumps({"fp64_sample": rows[-1]}), file=sys.stderr, flush=True)
        mean_s = float(np.mean([r["fp64_seconds"] for r in rows]))
        fp64 = {
            "reference": "plain FP64 march of the beamlet on the whole 200x200x300 phantom "
                         "(bitwise the CPU oracle)",
            "samples": rows,
            "fp64_beamlets_per_s": 1.0 / mean_s,
            "note": "beamlets at evenly spaced energy ranks of the batch, lowest to highest; "
                    "the FP64 beamlets/s is the inverse of the sample's mean time",
            "worst_max_dose_rel_err": max(abs(r["max_dose_rel_err"]) for r in rows),
            "min_gamma_pass_percent": min(r["gamma_1pct_1mm_pass_percent"] for r in rows),
        }

    value = B * K / elapsed
    cpu = None
    if not args.no_cpu_baseline:
        mid = int(np.argsort([b.energy for b in beamlets])[B // 2])
        cpu = batch_cpu_baseline(phantom, beamlets[mid], boxes[mid], cfg, args.cpu_sample_steps,
                                 float((box_cells * steps).sum()) / B)
    line = {
        "metric": "beamlet dose solves/s at <=0.5% max-dose err vs plain FP64; cell-step updates/s",
        "value": value,
        "unit": "beamlet dose solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk / f64 closure+projection+dose",
        "data": "synthetic seeded phantom and beamlets (no external data)",
        "config": {
            "workload": "256 pencil beamlets in one batch, 200x200x300 @1 mm water phantom with "
                        "12 bone/lung ellipsoids, 64x64xdepth boxes, Wentzel+nuclear, IF-SSPRK3, "
                        "FP32 bulk with tile skipping" +
                        (" and the tile-skip network" if use_net else ""),
            "n_beamlets": B,
            "tile_shape": list(handle.tile_shape),
            "launch": tuned["mode"] if tuned else "per_tile",
            "box_cells_total": int(box_cells.sum()),
            "energy_steps_per_beamlet": {"min": int(steps.min()), "mean": float(steps.mean()),
                                         "max": int(steps.max())},
            "tile_threshold": cfg.device.tile_threshold,
        },
        "cell_steps_per_s": cell_steps / elapsed,
        "cell_steps_per_s_evaluated": cell_steps * eval_frac / elapsed,
        "full_domain_equivalent_cell_steps_per_s": float(n_phantom * steps.sum()) * K / elapsed,
        "tiles_evaluated_fraction": eval_frac,
        "tile_skip_network": net,
        "autotune": tuned,
        "host_step_setup_ms_per_batch": host_ms / K,
        "setup_s": setup_s,
        "kernel_ms_per_step": {k: round(v[1] / K, 3) for k, v in stats.items() if v[0]},
        "fp64_full_domain": fp64,
        "speedup_vs_fp64_full_domain": (value / fp64["fp64_beamlets_per_s"]) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_combined(args):
    """The combined five-material workload, the part that can be marched:
    the beamlets whose entry footprint holds no air, in one batch with the
    double closure, the vacuum-aware CFL, the tile-skip network and graph
    steps; accuracy and the FP64 rate from the FP64 moving-domain march with
    the same CFL rule (the plain FP64 march is air-bound on this phantom)."""
    from proton_solver import _native, metrics, tileskip
    from proton_solver.batch import (air_free_beamlets, batch_workload_config, beamlet_boxes,
                                     combined_workload, full_problem, prepare_batch)

    phantom, beamlets = combined_workload(n_beamlets=args.beamlets)
    cfg = batch_workload_config(vacuum_aware_cfl=True)
    n = cfg.numerics
    all_boxes = beamlet_boxes(phantom, beamlets, n.energy_min, n.energy_window_sigma)
    idx = air_free_beamlets(phantom, beamlets, all_boxes)
    chosen = [beamlets[i] for i in idx]
    boxes = [all_boxes[i] for i in idx]
    handle, boxes = prepare_batch(phantom, chosen, cfg, boxes)
    num = _native.numerics_struct(cfg, _native.FP32_BULK_F64_CLOSURE)
    use_net = tileskip.WEIGHTS_FILE.exists() and args.skip_network != "off"
    tileskip.enable(handle, tileskip.load_model(0) if use_net else None, tileskip.SkipOptions())
    handle.set_graphs(True)
    for _ in range(args.warmup):
        handle.march(num, want_dose=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, steps = handle.march(num, want_dose=False)
    elapsed = time.perf_counter() - t0
    K = max(args.steps, 1)
    B = len(chosen)
    doses = [handle.beamlet_dose(b) for b in range(B)]
    # one profiled march (launched step by step) for the dominant kernel
    handle.set_graphs(False)
    handle.set_profiling(True)
    handle.reset_stats()
    handle.march(num, want_dose=False)
    stats = handle.kernel_stats()
    ev, full = handle.tile_stats()
    handle.set_profiling(False)
    tx, ty, tz = handle.tile_shape
    box_cells = np.array([b[2] * b[3] * b[4] for b in boxes], dtype=np.float64)
    tile_cap = np.array([-(-b[2] // tx) * -(-b[3] // ty) * -(-b[4] // tz) * (tx * ty * tz)
                         for b in boxes], dtype=np.float64)
    fill = box_cells.sum() / tile_cap.sum()
    roof = roofline_of(stats, STAGE_BYTES_FP32,
                       lambda launches: ev * float(tx * ty * tz) * fill / launches)

    print(json.dumps({"combined_batch": {"beamlets": B, "seconds": elapsed / K,
                                         "tiles_evaluated_fraction": ev / full if full else 1.0}}),
          file=sys.stderr, flush=True)

    # FP64 samples at evenly spaced energy ranks below 150 MeV: above it the
    # moving box grows with the scattered tails past the air-free footprint
    # and the FP64 march becomes air-bound.
    rows = []
    if args.fp64_sample > 0:
        low = [b for b in np.argsort([bl.energy for bl in chosen]) if chosen[b].energy < 150.0]
        picks = [int(low[int(round(q * (len(low) - 1)))])
                 for q in np.linspace(0.0, 1.0, args.fp64_sample)] if low else []
        ctx = _native.Context(0)
        num64 = _native.numerics_struct(cfg, _native.FP64)
        for b in picks:
            print(json.dumps({"fp64_moving_start": {"beamlet": idx[b],
                                                    "energy_MeV": round(chosen[b].energy, 3)}}),
                  file=sys.stderr, flush=True)
            ctx.set_problem(*full_problem(phantom, chosen[b]))
            ctx.set_state(None)
            t1 = time.perf_counter()
            E64, _, _ = ctx.march_moving(num64, n.active_guard_cells, n.active_truncation_sigma)
            sec = time.perf_counter() - t1
            full_dose = ctx.get_dose()
            x0, y0, nx, ny, nz = boxes[b]
            ref = full_dose[x0:x0 + nx, y0:y0 + ny, :nz]
            rows.append({
                "beamlet": idx[b], "energy_MeV": round(chosen[b].energy, 3),
                "fp64_moving_seconds": round(sec, 3), "fp64_energy_steps": len(E64) - 1,
                "batch_energy_steps": int(steps[b]),
                "max_dose_rel_err": float(doses[b].max() / full_dose.max() - 1.0),
                "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(
                    doses[b], ref, phantom.spacing, 1.0, 1.0, 10.0, device=0),
            })
            print(json.dumps({"fp64_moving_sample": rows[-1]}), file=sys.stderr, flush=True)
    value = B * K / elapsed
    fp64_rate = 1.0 / float(np.mean([r["fp64_moving_seconds"] for r in rows])) if rows else None
    line = {
        "metric": "beamlet dose solves/s at <=0.5% max-dose err vs plain FP64; cell-step updates/s",
        "value": value,
        "unit": "beamlet dose solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk / f64 closure+projection+dose",
        "data": "synthetic seeded phantom and beamlets (no external data)",
        "config": {
            "workload": f"combined: the {B} of {len(beamlets)} beamlets with no air within 48 mm "
                        "of their entry column, one batch over the 256x256x200 @1 mm phantom of "
                        "40 ellipsoids of the five materials, vacuum-aware CFL, graph steps" +
                        (", tile-skip network" if use_net else ""),
            "n_beamlets": B,
            "tile_shape": list(handle.tile_shape),
            "energy_steps_per_beamlet": {"min": int(steps.min()), "mean": float(steps.mean()),
                                         "max": int(steps.max())},
        },
        "cell_steps_per_s": float((box_cells * steps).sum()) * K / elapsed,
        "tiles_evaluated_fraction": ev / full if full else 1.0,
        "kernel_ms_per_march_profiled": {k: round(v[1], 3) for k, v in stats.items() if v[0]},
        "fp64_moving_domain": {
            "reference": "FP64 moving-domain march with the vacuum-aware CFL on the whole phantom "
                         "(samples below 150 MeV, where its box stays clear of air)",
            "samples": rows, "fp64_beamlets_per_s": fp64_rate,
        },
        "speedup_vs_fp64_moving_domain": (value / fp64_rate) if fp64_rate else None,
        "roofline": roof,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="batch256", choices=["batch256", "single", "combined"])
    ap.add_argument("--precision", default="fp32_bulk", choices=["fp32_bulk", "fp64", "fp64_fast"])
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="skip the FP64 solves used to state the error of the measured march")
    ap.add_argument("--cpu-sample-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--beamlets", type=int, default=256)
    ap.add_argument("--fp64-sample", type=int, default=3)
    ap.add_argument("--tile-threshold", type=float, default=None,
                    help="tile threshold relative to the peak inflow fluence "
                         "(default: proton_solver.batch.BATCH_TILE_THRESHOLD)")
    ap.add_argument("--skip-network", default="auto", choices=["auto", "on", "off"],
                    help="tile-skip network in the batched march (auto: when its weights exist)")
    ap.add_argument("--no-network-reference", action="store_true",
                    help="skip the untimed march without the network")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="default tile shape and launch instead of the autotuner's choice")
    ap.add_argument("--autotune-sample", type=int, default=32)
    ap.add_argument("--tile-shape", type=int, default=None,
                    help="index of the batched kernels' tile shape (no autotuning)")
    ap.add_argument("--profile-calibration", action="store_true",
                    help="run a kernel of known HBM traffic after the timed march")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path (beamlets are "
              "independent solves and would be sharded across GPUs by cost).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch256":
        args.steps = 1 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        # about 10 s of oracle work on the host's cores
        args.cpu_sample_steps = 100 if args.cpu_sample_steps is None else args.cpu_sample_steps
        if args.no_fp64_reference:
            args.fp64_sample = 0
        run_batch(args)
        return
    if args.workload == "combined":
        args.steps = 1 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        if args.beamlets == 256:
            args.beamlets = 1024
        if args.no_fp64_reference:
            args.fp64_sample = 0
        run_combined(args)
        return
    args.steps = 3 if args.steps is None else args.steps
    args.warmup = 1 if args.warmup is None else args.warmup
    args.cpu_sample_steps = 20 if args.cpu_sample_steps is None else args.cpu_sample_steps
    run_single(args)


def run_single(args):
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
    stage_bytes = STAGE_BYTES_FP32_SINGLE if precision is Precision.FP32_BULK else STAGE_BYTES_FP64

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
