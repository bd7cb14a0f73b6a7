"""Benchmark of the MI355X M1 energy march (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload batch256|single|interface|combined]

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

Workload "interface": one 200 MeV beam (sigma 3 mm) 2 mm from a water /
cortical-bone interface (bone fills x > 0), 160 x 160 x 480 at 0.5 mm, solved
by the batched march as one beamlet on the whole grid with its tile-skip
network; a step is one beam solve, against the plain FP64 march.

The whole command keeps to --time-budget seconds (default 450): the timed
and warm-up batches are cut to fit, and the counts actually run are the
reported steps and warmup.

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

T_START = time.perf_counter()
PEAK_HBM_GBS = 8000.0
MEV = 1.0

# Algorithmic HBM bytes per cell of each stage kernel: the compulsory traffic
# of a three-stage step with perfect halo reuse (each stage input read once,
# each output written once, 1-byte material id), split over the stages.
# FP64 (s = 8 B per value, dose read and written):
#   stage 0: read U (32) + ids (1), write V1 (32)
#   stage 1: read V1 (32) + U (32) + ids (1), write V2 (32)
#   stage 2: read V2 (32) + U (32) + V1.phi (8) + ids (1), write U (32), dose RMW (16)
STAGE_BYTES_FP64 = {"stage0": 65, "stage1": 97, "stage2": 121}
# FP32 bulk (s = 4) with a compensated FP64 dose pair (value and compensation
# read and written, 32 B):
#   stage 0: read U (16) + ids (1), write V1 (16)
#   stage 1: read V1 (16) + U (16) + ids (1), write V2 (16)
#   stage 2: read V2 (16) + U (16) + V1.phi (4) + ids (1), write U (16), dose pair RMW (32)
# The batch moves more than this (its stages 0 and 1 also store the next
# stage's projected stencil input, 32 B) and less (its compensated dose is a
# float pair, 16 B); the measured traffic shows what it moves.
STAGE_BYTES_FP32 = {"stage0": 33, "stage1": 49, "stage2": 85}
STAGE_BYTES_FP32_SINGLE = STAGE_BYTES_FP32
# Per cell-step (the three stages): 283 B in FP64, 167 B in FP32 bulk.
STEP_BYTES_FP64 = sum(STAGE_BYTES_FP64.values())
STEP_BYTES_FP32 = sum(STAGE_BYTES_FP32.values())
# HBM bytes per evaluated cell of the batched stage kernels per (tile shape,
# precision), from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the
# same counters on a kernel of known traffic (the summary named in
# TRAFFIC_PROFILE); shapes without a measurement report traffic null.
TRAFFIC_PROFILE = "profiles/stage_kernels_batch_pmc_traffic.md"
PMC_BYTES_PER_CELL_FP32 = {
    ((8, 16, 16), "fp32_bulk"): {"stage0": 101.8, "stage1": 151.3, "stage2": 153.3},
}
# Fraction of the VALU issue slots the batched stage kernels use
# (rocprofv3 SQ_INSTS_VALU x 4 cycles / (kernel time x 2.4 GHz x 1024
# SIMDs), VALU_PROFILE): the bound of these kernels.
VALU_PROFILE = "profiles/stage_kernels_batch_valu.md"
VALU_ISSUE_FRAC_FP32 = {
    ((8, 16, 16), "fp32_bulk"): {"stage0": 0.81, "stage1": 0.73, "stage2": 0.71},
}


def whole_step_roofline(cell_steps, seconds, bytes_per_cell_step):
    """HBM fraction of the whole march: cell-steps x algorithmic bytes per
    cell-step / seconds / peak (kernel launches, host work and idle time
    included)."""
    if not seconds > 0:
        return None
    gbs = cell_steps * bytes_per_cell_step / seconds / 1e9
    return {"bytes_per_cell_step": bytes_per_cell_step, "achieved": round(gbs, 2),
            "frac": round(gbs / PEAK_HBM_GBS, 5), "unit": "GB/s"}


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
    grid, mm, inflow = host_problem(cfg)
    return cpu_baseline_problem(grid, mm, inflow, cfg, n_steps_full, sample_steps)


def cpu_baseline_problem(grid, mm, inflow, cfg, n_steps_full, sample_steps):
    """cpu_baseline of a given problem (grid, material map, inflow)."""
    from oracle.oracle import OracleProblem

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
        "traffic_profile": TRAFFIC_PROFILE if traffic_per_cell else None,
        "valu_issue_frac_profiled": (valu_frac or {}).get(name),
        "valu_profile": VALU_PROFILE if valu_frac else None,
    }


def batch_config(tile_threshold=None, precision=None):
    from proton_solver.batch import BATCH_TILE_THRESHOLD, batch_workload_config
    from proton_solver.enums import Precision

    return batch_workload_config(BATCH_TILE_THRESHOLD if tile_threshold is None
                                 else tile_threshold,
                                 precision=Precision.FP32_BULK if precision is None else precision)


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


def plan_timed_steps(requested, batch_s, spent_s, budget_s, reserve_s, timed_cap_s):
    """Timed batches actually run: the requested count, cut so that the timed
    region lasts at most timed_cap_s and the whole command (spent_s so far,
    reserve_s still needed after the timed region) stays within budget_s;
    at least one."""
    room = min(float(timed_cap_s), float(budget_s) - float(spent_s) - float(reserve_s))
    fit = int(room // float(batch_s)) if batch_s > 0 else requested
    return max(1, min(int(requested), fit))


def plan_extra_warmup(requested, done, batch_s, spent_s, budget_s, reserve_s, timed_cap_s):
    """Whether one more untimed warm-up batch fits: at most two warm-ups, and
    only while the timed region it precedes keeps its full length."""
    if done >= min(int(requested), 2):
        return False
    timed = min(float(timed_cap_s), float(batch_s))
    return float(spent_s) + float(batch_s) + timed + float(reserve_s) <= float(budget_s)


def fp64_sample_order(energies, n):
    """Beamlet indices of the FP64 samples in the order they are run: the
    median energy rank first, then the lowest and the highest, then further
    evenly spaced ranks (the budget may stop the list early)."""
    B = len(energies)
    order = np.argsort(energies, kind="stable")
    ranks = [0.5, 0.0, 1.0] + [q for q in np.linspace(0.0, 1.0, max(n, 3) + 2)[1:-1]
                                if q not in (0.5,)]
    out = []
    for q in ranks:
        b = int(order[int(round(q * (B - 1)))])
        if b not in out:
            out.append(b)
        if len(out) == n:
            break
    return out


def fp64_batch_rate(samples, batch_steps):
    """FP64 beamlets/s of the whole batch from the sampled FP64 marches: the
    full-phantom march costs about the same per energy step whatever the
    beamlet, so the batch's FP64 time is the sampled seconds per step times
    the FP64 steps of every beamlet, estimated as its batch steps times the
    sampled FP64 / batch step ratio."""
    sec_per_step = float(np.mean([r["fp64_seconds"] / r["fp64_energy_steps"] for r in samples]))
    ratio = float(np.mean([r["fp64_energy_steps"] / r["batch_energy_steps"] for r in samples]))
    total = sec_per_step * ratio * float(np.sum(batch_steps))
    return len(batch_steps) / total, sec_per_step, ratio


DTYPE_LABEL = {
    "fp32_bulk": "f32 (state, WENO-Z, stage algebra, closure with error-free 1-f^2, HLL) / "
                 "f64 IF factors and inflow / compensated f32-pair dose",
    "fp32_bulk_f64_closure": "f32 bulk (state, WENO-Z, stage algebra) / f64 face projection, "
                             "closure, eigenvalues, HLL, IF factors / compensated f32-pair dose",
}


def run_batch(args):
    from proton_solver import Precision, _native, metrics, tileskip
    from proton_solver.batch import batch_workload, beamlet_boxes, full_problem, prepare_batch

    precision = {"fp32_bulk": Precision.FP32_BULK,
                 "fp32_bulk_f64_closure": Precision.FP32_BULK_F64_CLOSURE}[args.batch_precision]
    cfg = batch_config(args.tile_threshold, precision)
    phantom, beamlets = batch_workload(n_beamlets=args.beamlets)
    B = len(beamlets)

    def spent():
        return time.perf_counter() - T_START

    use_net = args.skip_network == "on" or (args.skip_network == "auto" and
                                            tileskip.WEIGHTS_FILE.exists())
    net_opts = tileskip.SkipOptions()
    model = tileskip.load_model(0) if use_net else None

    # Tile shape and launch from the autotuner (a sample of the beamlets,
    # outside the timed region), timed with the network the march uses.
    tuned = None
    if args.autotune and args.tile_shape is None:
        from proton_solver.autotune import autotune

        t0 = time.perf_counter()
        choice = autotune(phantom, beamlets, cfg, sample=args.autotune_sample,
                          modes=("per_tile", "device", "device_graphs"),
                          log=lambda row: print(json.dumps({"autotune": row}), file=sys.stderr,
                                                flush=True),
                          setup=lambda h: tileskip.enable(h, model, net_opts))
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
    else:
        from proton_solver.autotune import set_mode

        set_mode(handle, args.launch)
    setup_s = time.perf_counter() - t0
    num = _native.numerics_struct(cfg)
    tileskip.enable(handle, model, net_opts)

    # Untimed warm-up.  The first one is the profiled march: HIP events on
    # the march stream around every kernel (which launches every step one by
    # one) give the kernel times and the roofline; the timed batches run
    # without them.
    handle.set_profiling(True)
    handle.reset_stats()
    t0 = time.perf_counter()
    handle.march(num, want_dose=False)
    profiled_s = time.perf_counter() - t0
    stats = handle.kernel_stats()
    ev_p, full_p = handle.tile_stats()
    handle.set_profiling(False)
    print(json.dumps({"warmup": {"profiled_batch_s": round(profiled_s, 3),
                                 "spent_s": round(spent(), 1)}}), file=sys.stderr, flush=True)
    batch_s = profiled_s
    warm_done = 1
    reserve = (1.2 * batch_s if (use_net and not args.no_network_reference) else 0.0) + \
        (args.fp64_reserve_s if args.fp64_sample > 0 else 0.0) + \
        (args.cpu_reserve_s if not args.no_cpu_baseline else 0.0)
    while plan_extra_warmup(args.warmup, warm_done, batch_s, spent(), args.time_budget, reserve,
                            args.timed_seconds):
        t0 = time.perf_counter()
        handle.march(num, want_dose=False)
        batch_s = time.perf_counter() - t0
        warm_done += 1
    K = plan_timed_steps(args.steps, batch_s, spent(), args.time_budget, reserve,
                         args.timed_seconds)
    print(json.dumps({"plan": {"warmup": warm_done, "steps": K, "batch_s": round(batch_s, 3),
                               "reserve_s": round(reserve, 1), "spent_s": round(spent(), 1)}}),
          file=sys.stderr, flush=True)
    handle.reset_stats()
    host_ms0 = handle.host_ms()
    t0 = time.perf_counter()
    for _ in range(K):
        dose, steps = handle.march(num, want_dose=False)
    elapsed = time.perf_counter() - t0
    ev, full = handle.tile_stats()
    host_ms = handle.host_ms() - host_ms0
    ev_per_batch = ev / K

    # The network's effect: the same batch marched with the dilation rule alone
    # (outside the timed region).
    net = None
    beamlet_dose = handle.beamlet_dose
    if use_net:
        judged, rejections, back = handle.skip_stats()
        net_doses = [handle.beamlet_dose(b) for b in range(B)]
        beamlet_dose = net_doses.__getitem__
        net = {"weights": tileskip.WEIGHTS_FILE.name, "guard": net_opts.guard,
               "decision": net_opts.decision,
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
                "tiles_vs_dilation_only": ev_per_batch / ev0,
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
        return ev_p * float(tx * ty * tz) * fill / launches

    roof = roofline_of(stats, STAGE_BYTES_FP32, cells_per_launch,
                       PMC_BYTES_PER_CELL_FP32.get((tuple(handle.tile_shape), args.batch_precision)),
                       VALU_ISSUE_FRAC_FP32.get((tuple(handle.tile_shape), args.batch_precision)))
    if roof is not None:
        roof["measured_in"] = "the profiled warm-up batch (untimed)"
        roof["whole_step"] = whole_step_roofline(cell_steps * eval_frac, elapsed, STEP_BYTES_FP32)
    if args.profile_calibration:
        # a kernel of known traffic for scaling the HBM counters: the beamlet
        # dose read-out reads a float pair and writes a double per box cell
        x0, y0, nx, ny, nz = boxes[0]
        handle.beamlet_dose(0)
        print(json.dumps({"calibration_kernel": "box_dose_kernel", "cells": nx * ny * nz,
                          "bytes_read": 8 * nx * ny * nz, "bytes_written": 8 * nx * ny * nz}),
              file=sys.stderr, flush=True)

    # Plain FP64 full-domain march of a sample of the beamlets (outside the
    # timed region), as many as the budget allows.
    fp64 = None
    if args.fp64_sample > 0:
        picks = fp64_sample_order([b.energy for b in beamlets], args.fp64_sample)
        ctx = _native.Context(0)
        num64 = _native.numerics_struct(cfg, _native.FP64)
        rows = []
        skipped = []
        sec_per_step = args.fp64_prior_step_s
        for b in picks:
            predicted = sec_per_step * 1.15 * float(steps[b])
            left = args.time_budget - spent() - (args.cpu_reserve_s if not args.no_cpu_baseline
                                                 else 0.0)
            if rows and predicted > left:
                skipped.append({"beamlet": b, "energy_MeV": round(beamlets[b].energy, 3),
                                "predicted_s": round(predicted, 1)})
                continue
            grid, mm, inflow, beam = full_problem(phantom, beamlets[b])
            ctx.set_problem(grid, mm, inflow, beam)
            t1 = time.perf_counter()
            E64, _ = ctx.march(num64)
            sec = time.perf_counter() - t1
            ref = ctx.get_dose()
            sec_per_step = sec / (len(E64) - 1)
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
                    np.ascontiguousarray(got[sl]), np.ascontiguousarray(ref[sl]),
                    phantom.spacing, 1.0, 1.0, 10.0, device=0),
                "fp64_dose_outside_gamma_region_over_max": float(outside.max() / ref.max()),
            })
            print(json.dumps({"fp64_sample": rows[-1]}), file=sys.stderr, flush=True)
        rate, sps, ratio = fp64_batch_rate(rows, steps)
        fp64 = {
            "reference": "plain FP64 march of the beamlet on the whole 200x200x300 phantom "
                         "(bitwise the CPU oracle)",
            "samples": rows,
            "skipped_for_time": skipped,
            "fp64_seconds_per_energy_step": sps,
            "fp64_over_batch_energy_steps": ratio,
            "fp64_beamlets_per_s": rate,
            "fp64_beamlets_per_s_sample_mean": 1.0 / float(np.mean([r["fp64_seconds"]
                                                                     for r in rows])),
            "note": "fp64_beamlets_per_s is the whole batch at the sampled FP64 seconds per "
                    "energy step and FP64/batch step ratio; the samples are the median, "
                    "lowest and highest energy ranks, in that order, as the time budget allows",
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
        "steps": K,
        "warmup": warm_done,
        "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_LABEL[args.batch_precision],
        "data": "synthetic seeded phantom and beamlets (no external data)",
        "config": {
            "workload": "256 pencil beamlets in one batch, 200x200x300 @1 mm water phantom with "
                        "12 bone/lung ellipsoids, 64x64xdepth boxes, Wentzel+nuclear, IF-SSPRK3, "
                        "FP32 bulk with tile skipping" +
                        (" and the tile-skip network" if use_net else ""),
            "n_beamlets": B,
            "precision": args.batch_precision,
            "tile_shape": list(handle.tile_shape),
            "launch": tuned["mode"] if tuned else args.launch,
            "box_cells_total": int(box_cells.sum()),
            "energy_steps_per_beamlet": {"min": int(steps.min()), "mean": float(steps.mean()),
                                         "max": int(steps.max())},
            "tile_threshold": cfg.device.tile_threshold,
            "requested_steps": args.steps, "requested_warmup": args.warmup,
            "time_budget_s": args.time_budget,
        },
        "cell_steps_per_s": cell_steps / elapsed,
        "cell_steps_per_s_evaluated": cell_steps * eval_frac / elapsed,
        "full_domain_equivalent_cell_steps_per_s": float(n_phantom * steps.sum()) * K / elapsed,
        "tiles_evaluated_fraction": eval_frac,
        "tile_skip_network": net,
        "autotune": tuned,
        "host_step_setup_ms_per_batch": host_ms / K,
        "setup_s": setup_s,
        "kernel_ms_per_batch_profiled": {k: round(v[1], 3) for k, v in stats.items() if v[0]},
        "profiled_batch_s": profiled_s,
        "fp64_full_domain": fp64,
        "speedup_vs_fp64_full_domain": (value / fp64["fp64_beamlets_per_s"]) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "wall_s": round(spent(), 1),
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
    from proton_solver.autotune import set_mode

    set_mode(handle, "device_graphs")  # device-side stepping, graph-replayed steps
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
    set_mode(handle, "device")
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
    cpu = None
    if not args.no_cpu_baseline:
        mid = int(np.argsort([b.energy for b in chosen])[B // 2])
        cpu = batch_cpu_baseline(phantom, chosen[mid], boxes[mid], cfg,
                                 args.cpu_sample_steps or 100,
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
            "workload": f"combined, SUBSET: the {B} of {len(beamlets)} beamlets with no air within "
                        "48 mm of their entry column, one batch over the 256x256x200 @1 mm phantom "
                        "of 40 ellipsoids of the five materials, vacuum-aware CFL, graph steps" +
                        (", tile-skip network" if use_net else "") +
                        "; the other beamlets are air-bound (air's CFL step is 1/934 of water's) "
                        "and not marched",
            "n_beamlets": B,
            "n_beamlets_of_workload": len(beamlets),
            "coverage_fraction": B / len(beamlets),
            "tile_shape": list(handle.tile_shape),
            "energy_steps_per_beamlet": {"min": int(steps.min()), "mean": float(steps.mean()),
                                         "max": int(steps.max())},
        },
        "cell_steps_per_s": float((box_cells * steps).sum()) * K / elapsed,
        "tiles_evaluated_fraction": ev / full if full else 1.0,
        "kernel_ms_per_march_profiled": {k: round(v[1], 3) for k, v in stats.items() if v[0]},
        "fp64_moving_domain": {
            "reference": "FP64 moving-domain march with the vacuum-aware CFL on the whole phantom, "
                         "not the plain FP64 full-domain march (air-bound on this phantom)",
            "samples": rows, "fp64_beamlets_per_s": fp64_rate,
            "verified_energy_range_MeV": [min((r["energy_MeV"] for r in rows), default=None),
                                          max((r["energy_MeV"] for r in rows), default=None)],
            "unverified": "beamlets of 150 MeV and above: their FP64 moving box grows past the "
                          "air-free footprint and that march is air-bound too",
        },
        "speedup_vs_fp64_moving_domain": (value / fp64_rate) if fp64_rate else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def air_validation(args):
    """The quasi-steady air treatment against the plain FP64 march through a
    60 mm air pocket (semi-axes 12, 12, 30 mm at 15-75 mm depth in 48 x 48 x
    160 cells of water at 1 mm, 100 MeV), where the plain march takes
    air-bound energy steps: the check of the method the combined workload
    relies on, on a pocket as long along the beam as its longest (its own
    plain FP64 march would take hours per beamlet)."""
    import dataclasses

    from proton_solver import Material, _native, metrics
    from proton_solver.autotune import set_mode
    from proton_solver.batch import (Beamlet, Ellipsoid, batch_workload_config, ellipsoid_phantom,
                                     full_problem, prepare_batch)

    ph = ellipsoid_phantom((48, 48, 160), (0.001,) * 3,
                           [Ellipsoid((0.0, 0.0, 0.045), (0.012, 0.012, 0.030), Material.AIR)])
    beamlets = [Beamlet(energy=100.0, sigma=(0.003, 0.003), center=(0.0, 0.0))]
    cfg = batch_workload_config()
    cfg = dataclasses.replace(cfg, numerics=dataclasses.replace(
        cfg.numerics, max_energy_step=args.air_max_step))
    handle, _ = prepare_batch(ph, beamlets, cfg, boxes=[(0, 0, 48, 48, 160)], tile_shape=2)
    set_mode(handle, "device")
    handle.set_quasi_steady(1 << ph.materials.index(Material.AIR), args.qs_iters, args.qs_cfl)
    handle.set_quasi_steady_tolerance(args.qs_tol)
    t0 = time.perf_counter()
    _, steps = handle.march(_native.numerics_struct(cfg), want_dose=False)
    qs_s = time.perf_counter() - t0
    got = handle.beamlet_dose(0)
    handle.close()
    ctx = _native.Context(0)
    ctx.set_problem(*full_problem(ph, beamlets[0]))
    ctx.set_state(None)
    print(f"air validation: quasi-steady march {int(steps[0])} steps in {qs_s:.2f} s; "
          "plain FP64 march", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
    fp64_s = time.perf_counter() - t0
    ref = ctx.get_dose()
    return {
        "problem": "100 MeV beam through a 60 mm air pocket at 15-75 mm depth, 48x48x160 @1 mm "
                   "water",
        "reference": "plain FP64 march (air-bound steps; bitwise the CPU oracle)",
        "fp64_energy_steps": len(E64) - 1, "fp64_seconds": round(fp64_s, 2),
        "quasi_steady_energy_steps": int(steps[0]), "quasi_steady_seconds": round(qs_s, 2),
        "max_dose_rel_err": metrics.maximum_dose_relative_difference(got, ref),
        "max_abs_diff_over_max": metrics.max_abs_difference_over_max(got, ref),
        "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(got, ref, ph.spacing, 1.0, 1.0,
                                                               10.0, device=0),
    }


def combined_sample_picks(phantom, beamlets, boxes, air_free, n):
    """Beamlets of the combined workload checked against the FP64
    moving-domain march: n energy quantiles of the air-free beamlets that
    stop inside the phantom (box shallower than it).  Past the far face the
    moving domain of an exiting beam widens by ~50 cells and reaches air (a
    227 MeV beamlet's FP64 march drops to ~4e-4 MeV steps at 85 MeV there),
    hours per beamlet."""
    stopping = [b for b in air_free if boxes[b][4] < phantom.shape[2]]
    if n <= 0 or not stopping:
        return []
    order = sorted(stopping, key=lambda b: beamlets[b].energy)
    return sorted({order[int(round(q * (len(order) - 1)))] for q in np.linspace(0.0, 1.0, n)})


def run_combined_quasi_steady(args):
    """The combined five-material workload in full: all its beamlets in one
    batch with the air quasi-steady (held through each stage and relaxed
    towards div F = 0 after it), so that no beamlet takes air-bound steps;
    device-side stepping, the whole-phantom CFL bound over the other
    materials, a 0.1 MeV maximum step.  Accuracy: the method against the plain
    FP64 march through an air pocket (air_validation), and air-free beamlets
    of the batch against the FP64 moving-domain march at the same maximum
    step (the plain FP64 march of this phantom is air-bound: ~1e-3 MeV steps,
    hours per beamlet).  Not bound by --time-budget: two marches of the whole
    batch and the checks take several minutes."""
    import dataclasses

    from proton_solver import Material, _native, metrics
    from proton_solver.autotune import set_mode
    from proton_solver.batch import (air_free_beamlets, batch_workload_config, beamlet_boxes,
                                     combined_workload, full_problem, prepare_batch)

    def spent():
        return time.perf_counter() - T_START

    phantom, beamlets = combined_workload(n_beamlets=args.beamlets)
    cfg = batch_workload_config()
    cfg = dataclasses.replace(cfg, numerics=dataclasses.replace(
        cfg.numerics, max_energy_step=args.air_max_step))
    n = cfg.numerics
    boxes = beamlet_boxes(phantom, beamlets, n.energy_min, n.energy_window_sigma)
    air = phantom.materials.index(Material.AIR)
    air_free = air_free_beamlets(phantom, beamlets, boxes)
    handle, boxes = prepare_batch(phantom, beamlets, cfg, boxes, tile_shape=2)
    set_mode(handle, "device")
    handle.set_quasi_steady(1 << air, args.qs_iters, args.qs_cfl)
    handle.set_quasi_steady_tolerance(args.qs_tol)
    handle.set_progress(250)
    num = _native.numerics_struct(cfg)
    # profiled march (untimed): kernel times and the roofline
    handle.set_profiling(True)
    handle.reset_stats()
    t0 = time.perf_counter()
    handle.march(num, want_dose=False)
    profiled_s = time.perf_counter() - t0
    stats = handle.kernel_stats()
    ev_p, _ = handle.tile_stats()
    handle.set_profiling(False)
    print(json.dumps({"combined_profiled_batch_s": round(profiled_s, 2),
                      "spent_s": round(spent(), 1)}), file=sys.stderr, flush=True)
    for _ in range(max(args.warmup - 1, 0)):
        handle.march(num, want_dose=False)
    handle.reset_stats()
    K = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(K):
        _, steps = handle.march(num, want_dose=False)
    elapsed = time.perf_counter() - t0
    ev, full = handle.tile_stats()
    B = len(beamlets)
    tx, ty, tz = handle.tile_shape
    box_cells = np.array([b[2] * b[3] * b[4] for b in boxes], dtype=np.float64)
    tile_cap = np.array([-(-b[2] // tx) * -(-b[3] // ty) * -(-b[4] // tz) * (tx * ty * tz)
                         for b in boxes], dtype=np.float64)
    fill = box_cells.sum() / tile_cap.sum()
    roof = roofline_of(stats, STAGE_BYTES_FP32,
                       lambda launches: ev_p * float(tx * ty * tz) * fill / launches)

    # FP64 moving-domain samples of air-free beamlets, at the batch's maximum step
    rows = []
    picks = combined_sample_picks(phantom, beamlets, boxes, air_free, args.fp64_sample)
    if picks:
        cfg64 = dataclasses.replace(cfg, numerics=dataclasses.replace(cfg.numerics,
                                                                      vacuum_aware_cfl=True))
        num64 = _native.numerics_struct(cfg64, _native.FP64)
        ctx = _native.Context(0)
        for b in picks:  # about a second each (air-free, stopping inside the phantom)
            print(f"fp64 moving-domain march of beamlet {b} ({beamlets[b].energy:.1f} MeV)",
                  file=sys.stderr, flush=True)
            ctx.set_problem(*full_problem(phantom, beamlets[b]))
            ctx.set_state(None)
            t1 = time.perf_counter()
            E64, _, _ = ctx.march_moving(num64, n.active_guard_cells, n.active_truncation_sigma)
            sec = time.perf_counter() - t1
            full_dose = ctx.get_dose()
            x0, y0, nx, ny, nz = boxes[b]
            ref = full_dose[x0:x0 + nx, y0:y0 + ny, :nz]
            got = handle.beamlet_dose(b)
            rows.append({
                "beamlet": b, "energy_MeV": round(beamlets[b].energy, 3),
                "fp64_moving_seconds": round(sec, 3), "fp64_energy_steps": len(E64) - 1,
                "batch_energy_steps": int(steps[b]),
                "max_dose_rel_err": float(got.max() / full_dose.max() - 1.0),
                "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(
                    got, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0),
            })
            print(json.dumps({"fp64_moving_sample": rows[-1]}), file=sys.stderr, flush=True)
    valid = None
    if not args.no_fp64_reference:
        valid = air_validation(args)
        print(json.dumps({"air_validation": valid}), file=sys.stderr, flush=True)
    value = B * K / elapsed
    cpu = None
    if not args.no_cpu_baseline:
        mid = int(np.argsort([b.energy for b in beamlets])[B // 2])
        cpu = batch_cpu_baseline(phantom, beamlets[mid], boxes[mid], cfg,
                                 args.cpu_sample_steps or 100,
                                 float((box_cells * steps).sum()) / B)
    air_boxes = int(sum(1 for b in boxes if (phantom.ids[b[0]:b[0] + b[2], b[1]:b[1] + b[3],
                                                         :b[4]] == air).any()))
    line = {
        "metric": "beamlet dose solves/s at <=0.5% max-dose err vs plain FP64; cell-step updates/s",
        "value": value,
        "unit": "beamlet dose solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": max(args.warmup, 1),
        "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_LABEL["fp32_bulk"],
        "data": "synthetic seeded phantom and beamlets (no external data)",
        "config": {
            "workload": f"combined: all {B} beamlets in one batch over the 256x256x200 @1 mm "
                        "phantom of 40 ellipsoids of the five materials, air quasi-steady, "
                        "device-side steps, whole-phantom CFL bound over the other materials, "
                        f"max energy step {args.air_max_step} MeV",
            "n_beamlets": B,
            "boxes_holding_air": air_boxes,
            "air_free_beamlets": len(air_free),
            "quasi_steady": {"material": "AIR", "max_sweeps_per_stage": args.qs_iters,
                             "tolerance": args.qs_tol, "pseudo_cfl": args.qs_cfl},
            "tile_shape": list(handle.tile_shape),
            "energy_steps_per_beamlet": {"min": int(steps.min()), "mean": float(steps.mean()),
                                         "max": int(steps.max())},
        },
        "cell_steps_per_s": float((box_cells * steps).sum()) * K / elapsed,
        "tiles_evaluated_fraction": ev / full if full else 1.0,
        "kernel_ms_per_march_profiled": {k: round(v[1], 3) for k, v in stats.items() if v[0]},
        "air_validation": valid,
        "fp64_moving_domain_air_free": {
            "reference": "FP64 moving-domain march with the vacuum-aware CFL, same maximum step, "
                         "on air-free beamlets of the batch",
            "samples": rows,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "wall_s": round(spent(), 1),
    }
    print(json.dumps(line), flush=True)


def run_interface(args):
    """The lateral water / cortical-bone interface beam (160 x 160 x 480 at
    0.5 mm, 200 MeV) solved by the batched march as one beamlet on the whole
    grid with its tile-skip network; a step is one beam solve."""
    from proton_solver import _native, metrics, tileskip
    from proton_solver.batch import (batch_workload_config, full_problem, interface_workload,
                                     prepare_batch)

    cfg = batch_workload_config()
    num = _native.numerics_struct(cfg)
    phantom, beamlets, boxes = interface_workload(0.0)
    handle, _ = prepare_batch(phantom, beamlets, cfg, boxes=boxes)
    opts = tileskip.SkipOptions(judge_every=256)
    use_net = args.skip_network != "off" and tileskip.INTERFACE_WEIGHTS_FILE.exists()
    model = tileskip.load_model(0, tileskip.INTERFACE_WEIGHTS_FILE) if use_net else None
    n_cells = int(np.prod(phantom.shape))
    # the mask alone (untimed), for the network's effect
    tileskip.enable(handle, None)
    handle.reset_stats()
    t0 = time.perf_counter()
    _, steps_mask = handle.march(num, want_dose=False)
    mask_s = time.perf_counter() - t0
    mask_tiles, full = handle.tile_stats()
    dose_mask = handle.beamlet_dose(0)
    tileskip.enable(handle, model, opts)
    handle.set_profiling(True)
    handle.reset_stats()
    handle.march(num, want_dose=False)  # profiled warm-up
    stats = handle.kernel_stats()
    ev_p, _ = handle.tile_stats()
    handle.set_profiling(False)
    for _ in range(max(args.warmup - 1, 0)):
        handle.march(num, want_dose=False)
    handle.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, steps = handle.march(num, want_dose=False)
    elapsed = time.perf_counter() - t0
    ev, full = handle.tile_stats()
    dose = handle.beamlet_dose(0)
    K = max(args.steps, 1)
    tx, ty, tz = handle.tile_shape
    roof = roofline_of(stats, STAGE_BYTES_FP32,
                       lambda launches: ev_p * float(tx * ty * tz) / launches)
    eval_frac = ev / full if full else 1.0
    cell_steps = float(n_cells) * float(steps[0]) * K
    if roof is not None:
        roof["measured_in"] = "the profiled warm-up solve (untimed)"
        roof["whole_step"] = whole_step_roofline(cell_steps * eval_frac, elapsed, STEP_BYTES_FP32)
    accuracy = None
    if not args.no_fp64_reference:
        ctx = _native.Context(0)
        ctx.set_problem(*full_problem(phantom, beamlets[0]))
        ctx.set_state(None)
        t0 = time.perf_counter()
        E64, _ = ctx.march(_native.numerics_struct(cfg, _native.FP64))
        sec64 = time.perf_counter() - t0
        ref = ctx.get_dose()
        accuracy = {
            "reference": "plain FP64 march on the same grid (bitwise the CPU oracle)",
            "fp64_solve_s": sec64, "fp64_energy_steps": len(E64) - 1,
            "speedup_vs_fp64_march": sec64 / (elapsed / K),
        }
        for label, d in (("network", dose), ("mask_alone", dose_mask)):
            accuracy[label] = {
                "max_dose_rel_err": metrics.maximum_dose_relative_difference(d, ref),
                "max_abs_diff_over_max": metrics.max_abs_difference_over_max(d, ref),
                "gamma_1pct_1mm_pass_percent": metrics.gamma_pass_rate(
                    d, ref, phantom.spacing, 1.0, 1.0, 10.0, device=0)}
    cpu = None
    if not args.no_cpu_baseline:
        from proton_solver.batch import box_problem

        grid, mm, inflow, beam = box_problem(phantom, boxes[0], beamlets[0])
        cfg1 = type(cfg)(geometry=cfg.geometry, beam=beam, physics=cfg.physics,
                         numerics=cfg.numerics, device=cfg.device)
        cpu = cpu_baseline_problem(grid, mm, inflow, cfg1, int(steps[0]), args.cpu_sample_steps)
    line = {
        "metric": "beamlet dose solves/s at <=0.5% max-dose err vs plain FP64; cell-step updates/s",
        "value": K / elapsed,
        "unit": "beamlet dose solves/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE_LABEL["fp32_bulk"],
        "data": "synthetic half-space phantom (no external data)",
        "config": {
            "workload": "200 MeV beam 2 mm from a water / cortical-bone (x > 0) interface, "
                        "160x160x480 @0.5 mm, Wentzel+nuclear, IF-SSPRK3, FP32 bulk with tile "
                        "skipping" + (" and the tile-skip network" if use_net else ""),
            "energy_steps": int(steps[0]), "tile_shape": list(handle.tile_shape),
        },
        "cell_steps_per_s": cell_steps / elapsed,
        "cell_steps_per_s_evaluated": cell_steps * eval_frac / elapsed,
        "tiles_evaluated_fraction": eval_frac,
        "tile_skip_network": {"weights": tileskip.INTERFACE_WEIGHTS_FILE.name, "guard": opts.guard,
                              "judge_every": opts.judge_every,
                              "mask_alone_seconds": mask_s,
                              "tiles_vs_mask_alone": (ev / K) / mask_tiles if mask_tiles else None,
                              "max_dose_change_vs_mask_over_max":
                                  float(np.abs(dose - dose_mask).max() / dose_mask.max())}
        if use_net else None,
        "kernel_ms_per_solve_profiled": {k: round(v[1], 3) for k, v in stats.items() if v[0]},
        "accuracy_vs_fp64": accuracy,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="batch256",
                    choices=["batch256", "single", "combined", "interface"])
    ap.add_argument("--precision", default="fp32_bulk", choices=["fp32_bulk", "fp64", "fp64_fast"])
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="skip the FP64 solves used to state the error of the measured march")
    ap.add_argument("--cpu-sample-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--beamlets", type=int, default=256)
    ap.add_argument("--fp64-sample", type=int, default=3)
    ap.add_argument("--batch-precision", default="fp32_bulk",
                    choices=["fp32_bulk", "fp32_bulk_f64_closure"])
    ap.add_argument("--time-budget", type=float, default=450.0,
                    help="wall-clock seconds the whole command should stay within; the timed "
                         "and warm-up batches are cut to fit (the counts run are reported)")
    ap.add_argument("--timed-seconds", type=float, default=60.0,
                    help="upper bound of the timed region, seconds")
    ap.add_argument("--fp64-reserve-s", type=float, default=60.0,
                    help="seconds kept for the FP64 samples when planning the timed batches")
    ap.add_argument("--fp64-prior-step-s", type=float, default=0.03,
                    help="FP64 seconds per energy step assumed before the first sample")
    ap.add_argument("--cpu-reserve-s", type=float, default=30.0,
                    help="seconds kept for the CPU baseline")
    ap.add_argument("--tile-threshold", type=float, default=None,
                    help="tile threshold relative to the peak inflow fluence "
                         "(default: proton_solver.batch.BATCH_TILE_THRESHOLD)")
    ap.add_argument("--skip-network", default="auto", choices=["auto", "on", "off"],
                    help="tile-skip network in the batched march (auto: when its weights exist)")
    ap.add_argument("--no-network-reference", action="store_true",
                    help="skip the untimed march without the network")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="default tile shape and launch instead of the autotuner's choice")
    ap.add_argument("--autotune-sample", type=int, default=24)
    ap.add_argument("--air-free-only", action="store_true",
                    help="combined workload: march only the beamlets with no air near their "
                         "entry column (vacuum-aware CFL, network, graph steps) instead of all "
                         "of them with the air quasi-steady")
    ap.add_argument("--qs-iters", type=int, default=128,
                    help="quasi-steady air: at most this many relaxation sweeps after each stage")
    ap.add_argument("--qs-tol", type=float, default=1e-5,
                    help="quasi-steady air: a beamlet stops sweeping once no air cell changes "
                         "by more than this x its peak inflow fluence (0: every sweep)")
    ap.add_argument("--qs-cfl", type=float, default=0.3,
                    help="quasi-steady air: pseudo-time CFL number of a sweep")
    ap.add_argument("--air-max-step", type=float, default=0.1,
                    help="combined workload: maximum energy step, MeV")
    ap.add_argument("--launch", default="device_graphs",
                    choices=["per_tile", "persistent", "graphs", "device", "device_persistent",
                             "device_graphs"],
                    help="launch of the batch when the autotuner does not choose it: "
                         "host- or device-stepped energy steps, one workgroup per tile, a "
                         "persistent grid or graph replay")
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
        args.steps = 3 if args.steps is None else args.steps
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
        if args.air_free_only:
            run_combined(args)
        else:
            run_combined_quasi_steady(args)
        return
    args.steps = 3 if args.steps is None else args.steps
    args.warmup = 1 if args.warmup is None else args.warmup
    args.cpu_sample_steps = 20 if args.cpu_sample_steps is None else args.cpu_sample_steps
    if args.workload == "interface":
        run_interface(args)
        return
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

    # Untimed warm-up; the first warm-up march is the profiled one (HIP events
    # around every kernel), the timed marches run without the events.
    ctx.set_profiling(True)
    ctx.reset_stats()
    for w in range(max(args.warmup, 1)):
        ctx.set_state(None)
        ctx.march(num)
        if w == 0:
            stats = ctx.kernel_stats()
            tiles_eval, tiles_full = ctx.tile_stats()
            ctx.set_profiling(False)
    steps_per_march = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.set_state(None)
        E, _ = ctx.march(num)
        steps_per_march.append(len(E) - 1)
    elapsed = time.perf_counter() - t0
    dose = ctx.get_dose()

    n_energy_steps = steps_per_march[-1]
    ms_per_step = 1000.0 * elapsed / args.steps
    solves_per_s = args.steps / elapsed
    cell_steps = n_cells * sum(steps_per_march)

    # Dominant stage kernel and its HBM roofline (from the profiled warm-up
    # march).  Cells per launch are the cells of the evaluated tiles (all
    # cells when no tile is skipped).
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
            "measured_in": "the profiled warm-up march (untimed)",
            "whole_step": whole_step_roofline(
                cell_steps * eval_frac, elapsed,
                STEP_BYTES_FP32 if precision is Precision.FP32_BULK else STEP_BYTES_FP64),
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
