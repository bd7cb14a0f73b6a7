"""Tile-shape and launch autotuner for the batched march (no reference counterpart).

The batched kernels are compiled for several workgroup tile shapes
(`_native.batch_tile_shapes()`, from M1_BATCH_SHAPES in csrc/m1_device.h) and
can be launched three ways: one workgroup per tile, a persistent grid pulling
tiles from per-XCD queues, or plain steps replayed from HIP graphs; and the
energy steps can be set up on the host (one round trip per step) or on the
device ("device_*" modes: per-beamlet CFL steps, step scalars and judge
decisions in kernels, steps enqueued without waiting; the per-tile launch is
then sized from an estimate of the list length).  The tile
shape sets the skip granularity (finer tiles skip more vacuum but re-read
more halo cells per cell) and the register and LDS footprint of the stage
kernels, so the best one depends on the beamlets and the phantom.

``autotune`` marches a sample of the beamlets (evenly spaced energy ranks,
so the step counts and box depths of the whole batch are represented) with
every tile shape in the default launch, then the fastest shape in the other
launches, and returns the fastest combination with the table of timings.
Every candidate computes the same march up to the tile set (the tile shape
changes which cells are evaluated, not the arithmetic of those that are).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native

MODES = ("per_tile", "persistent", "graphs", "device", "device_persistent", "device_graphs")


@dataclass(frozen=True)
class Choice:
    shape_index: int
    shape: tuple[int, int, int]
    mode: str
    seconds: float
    table: list = field(default_factory=list, compare=False)

    def apply(self, handle: _native.Batch):
        """Set the launch of a batch prepared with tile shape `shape_index`."""
        set_mode(handle, self.mode)


def set_mode(handle: _native.Batch, mode: str):
    if mode not in MODES:
        raise ValueError(f"mode must be one of {MODES}")
    handle.set_device_steps(mode.startswith("device"))
    handle.set_persistent(mode.endswith("persistent"))
    handle.set_graphs(mode.endswith("graphs"))


def sample_beamlets(beamlets, n):
    """n beamlets at evenly spaced energy ranks (lowest and highest included)."""
    n = max(1, min(int(n), len(beamlets)))
    order = np.argsort([b.energy for b in beamlets], kind="stable")
    picks = sorted({int(order[int(round(q * (len(beamlets) - 1)))])
                    for q in np.linspace(0.0, 1.0, n)})
    return [beamlets[i] for i in picks]


# Graph replay is taken when it is within this fraction of the fastest launch:
# the sample's timings scatter by 1-2 % between runs, and on the whole
# 256-beamlet batch graph replay and the launched steps are within 0.3 %.
PREFER_WITHIN = 0.02


def autotune(phantom, beamlets, config, sample=12, shapes=None, modes=MODES, device=None,
             log=None, setup=None, prefer=("device_graphs",), prefer_within=PREFER_WITHIN):
    """Fastest (tile shape, launch) on a sample of the batch; returns a Choice
    whose table lists every timed candidate.  setup(handle), when given, is
    applied to every candidate batch before it is timed (e.g. the tile-skip
    network the timed march uses).  A launch in `prefer` is taken when it is
    within `prefer_within` (relative) of the fastest: graph replay submits
    one graph per energy step instead of a dozen launches, which leaves the
    host thread free, at a cost within the timing noise."""
    from .batch import prepare_batch

    sub = sample_beamlets(beamlets, sample)
    all_shapes = _native.batch_tile_shapes()
    idx = list(range(len(all_shapes))) if shapes is None else list(shapes)
    num = _native.numerics_struct(config)
    table = []

    def timed(handle, mode):
        set_mode(handle, mode)
        t0 = time.perf_counter()
        handle.march(num, want_dose=False)
        return time.perf_counter() - t0

    best = None
    handles = {}
    for n_i, i in enumerate(idx):
        handle, _ = prepare_batch(phantom, sub, config, device=device, tile_shape=i)
        if setup is not None:
            setup(handle)
        if n_i == 0:
            timed(handle, "per_tile")  # first launches of the process
        sec = timed(handle, "per_tile")
        table.append({"shape": all_shapes[i], "mode": "per_tile", "seconds": round(sec, 4)})
        if log:
            log(table[-1])
        handles[i] = handle
        if best is None or sec < best[2]:
            best = (i, "per_tile", sec)
    shape_i = best[0]
    for mode in modes:
        if mode == "per_tile":
            continue
        sec = timed(handles[shape_i], mode)
        table.append({"shape": all_shapes[shape_i], "mode": mode, "seconds": round(sec, 4)})
        if log:
            log(table[-1])
        if sec < best[2]:
            best = (shape_i, mode, sec)
    for row in table:
        if (row["mode"] in prefer and tuple(row["shape"]) == all_shapes[shape_i] and
                row["seconds"] <= best[2] * (1.0 + prefer_within)):
            best = (shape_i, row["mode"], row["seconds"])
    for h in handles.values():
        h.close()
    return Choice(best[0], all_shapes[best[0]], best[1], best[2], table)
