"""Learned tile skipping for the batched march (no reference counterpart).

The batched march evaluates a workgroup tile when it holds phi above the
tile threshold or when a cell within 9 cells of it does (the reach of one
three-stage step): the dilated conservative mask.  The network's candidates
are the tiles that mask adds and that hold nothing above the threshold
themselves; the tile-skip network (11 features per tile, three hidden
layers of 64 ReLU units, bf16 weights, float32 accumulation; the device
kernels are in csrc/m1_mlp.hip) predicts which candidates stay at or below
the threshold over the next `horizon` steps, and the march leaves those out
while they remain candidates (checked every step).  Tiles above the
threshold and inflow tiles are always evaluated.  (SkipOptions.guard = 4
restricts the candidates to tiles 4 cells away from anything above the
threshold, the earlier fixed rule.)  A periodic judge takes a beamlet's step
both ways from the same state; a beamlet whose step dose moves by more than
`judge_tol` of the step's largest increment is sent back to its last
accepted checkpoint and marches on with the mask alone (include/m1march.h,
m1_batch_set_skip_model).

Training data comes from the march itself: ``record_samples`` runs the
batched march with every tile of the active boxes evaluated and labels each
candidate tile by whether it stayed at or below the threshold over the next
`horizon` steps; ``against_guard_rule`` compares a trained network with the
fixed 4-cell guard rule on held-out samples.  ``train`` fits the network on
the device.  ``main`` records on a training workload -- the 256-beamlet
workload's construction with other seeds, or the bone-interface beam on its
whole grid at seeded beam offsets -- trains and writes the weights that
``load_model`` reads.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native

WEIGHTS_FILE = Path(__file__).with_name("tileskip_weights.npy")
INTERFACE_WEIGHTS_FILE = Path(__file__).with_name("tileskip_weights_interface.npy")

# Feature order of the device kernel (include/m1march.h).
FEATURES = ("log10_max_phi", "log10_mean_phi", "flux_ratio", "density", "energy",
            "nbr_mx", "nbr_px", "nbr_my", "nbr_py", "nbr_mz", "nbr_pz")


@dataclass(frozen=True)
class SkipOptions:
    decision: float = 0.9      # probability of staying empty needed to leave a tile out
    horizon: int = 8           # steps a tile is left out for (the label's window)
    infer_every: int = 4       # steps between evaluations of the network
    judge_every: int = 64      # steps between judged steps of a beamlet
    judge_tol: float = 1e-3    # allowed step-dose change over the step's largest increment
    guard: int = 0             # candidates: 0 every tile the reach rule adds, 4 only those
                               # 4 cells from any cell above the threshold


def record_samples(handle: _native.Batch, num, horizon=8, capacity=1 << 22,
                   full_domain=True, guard=0):
    """Run one march of a prepared batch with recording on; returns x (n, 11)
    float32 features, y (n,) float32 labels (1: the tile stayed empty) and
    near (n,) bool (a cell within 4 cells of the tile was above the threshold
    when the sample was taken; only with guard 0)."""
    handle.set_skip_model(None)
    handle.set_skip_guard(guard)
    handle.record(True, full_domain=full_domain, horizon=horizon, capacity=capacity)
    try:
        handle.march(num, want_dose=False)
        x, code = handle.samples()
    finally:
        handle.record(False)
    return x, np.mod(code, 2.0).astype(np.float32), code >= 2.0


def init_params(seed=0):
    """He-initialised float32 parameters, biases zero."""
    rng = np.random.default_rng(seed)
    n_in, n_h = len(FEATURES), 64
    parts = []
    for fan_in, shape in ((n_in, (n_in, n_h)), (None, (n_h,)), (n_h, (n_h, n_h)), (None, (n_h,)),
                          (n_h, (n_h, n_h)), (None, (n_h,)), (n_h, (n_h,)), (None, (1,))):
        if fan_in is None:
            parts.append(np.zeros(int(np.prod(shape))))
        else:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), size=shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def train(x, y, device=0, epochs=20, batch=4096, lr=1e-3, seed=0, params=None):
    """Fit a network on the device (Adam on the mean binary cross-entropy).
    Returns (model, final loss)."""
    model = _native.Mlp(device, init_params(seed) if params is None else params)
    loss = model.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return model, loss


def confusion(model: _native.Mlp, x, y, decision):
    """Counts on (x, y): tiles left out that did not stay empty (false_skip)
    and tiles left out that did (true_skip)."""
    p = model.forward(x)
    skip = p > decision
    empty = y > 0.5
    return {"samples": int(len(y)), "empty_fraction": float(empty.mean()) if len(y) else 0.0,
            "true_skip": int((skip & empty).sum()), "false_skip": int((skip & ~empty).sum()),
            "kept": int((~skip).sum())}


def against_guard_rule(model: _native.Mlp, x, y, near, decisions=(0.5, 0.7, 0.9, 0.95, 0.99)):
    """The network against the fixed 4-cell guard rule on candidates of every
    distance (guard 0 samples): per rule, the fraction of candidates left out
    and the fraction left out that did not stay empty (false skips, the
    fluence a skip loses).  The guard rule leaves out exactly the candidates
    with nothing above the threshold within 4 cells."""
    p = model.forward(x)
    empty = y > 0.5
    n = max(len(y), 1)

    def row(name, skip):
        k = int(skip.sum())
        return {"rule": name, "skip_fraction": k / n,
                "false_skip_fraction": float((skip & ~empty).sum()) / n,
                "false_skips_of_skipped": float((skip & ~empty).sum()) / max(k, 1)}

    rows = [row("guard 4 cells", ~near), row("dilation only", np.zeros(len(y), bool))]
    rows += [row(f"network p > {d}", p > d) for d in decisions]
    return {"samples": int(len(y)), "empty_fraction": float(empty.mean()) if len(y) else 0.0,
            "near_fraction": float(near.mean()) if len(y) else 0.0, "rules": rows}


def save_params(params, path=WEIGHTS_FILE):
    p = np.ascontiguousarray(params, dtype=np.float32)
    np.save(Path(path), p, allow_pickle=False)


def load_params(path=WEIGHTS_FILE):
    p = np.load(Path(path), allow_pickle=False)
    n = _native.load_library().m1_mlp_param_count()
    if p.dtype != np.float32 or p.shape != (n,):
        raise ValueError(f"{path}: expected {n} float32 parameters, got {p.dtype} {p.shape}")
    return p


def load_model(device=0, path=WEIGHTS_FILE) -> _native.Mlp:
    """The network with the weights written by ``main`` (the batched
    workload's; INTERFACE_WEIGHTS_FILE holds the interface workload's)."""
    if not Path(path).exists():
        raise FileNotFoundError(f"{path}: no tile-skip weights; run "
                                "python -m proton_solver.tileskip to train them")
    return _native.Mlp(device, load_params(path))


def enable(handle: _native.Batch, model, options: SkipOptions = SkipOptions()):
    """Use `model` (None: switch the network off) in the marches of `handle`."""
    handle.set_skip_guard(options.guard)
    handle.set_skip_model(model, options.decision, options.horizon, options.infer_every,
                          options.judge_every, options.judge_tol)


def training_workload(n_beamlets=96):
    """Phantom and beamlets of the training runs: the 256-beamlet workload's
    construction with other seeds (phantom 11, beamlets 110)."""
    from .batch import batch_workload

    return batch_workload(n_beamlets=n_beamlets, phantom_seed=11, beamlet_seed=110)


def record_interface_samples(offsets, guard=0, horizon=8, device=None):
    """Samples of the interface workload (one beam on the whole 160 x 160 x
    480 grid) from full-domain runs at the given beam offsets: every tile is
    evaluated and each candidate tile is labelled by whether it stayed at or
    below the threshold over the next `horizon` steps."""
    from .batch import batch_workload_config, interface_workload, prepare_batch

    cfg = batch_workload_config()
    num = _native.numerics_struct(cfg)
    xs, ys, nears = [], [], []
    for off in offsets:
        phantom, beamlets, boxes = interface_workload(off)
        handle, _ = prepare_batch(phantom, beamlets, cfg, boxes=boxes, device=device)
        try:
            x, y, near = record_samples(handle, num, horizon=horizon, guard=guard)
        finally:
            handle.close()
        xs.append(x)
        ys.append(y)
        nears.append(near)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(nears)


def main(argv=None):
    from .batch import batch_workload_config, interface_offsets, prepare_batch

    ap = argparse.ArgumentParser(description="record samples, train the tile-skip network "
                                             "and write its weights")
    ap.add_argument("--beamlets", type=int, default=96)
    ap.add_argument("--horizon", type=int, default=SkipOptions.horizon)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--decision", type=float, default=SkipOptions.decision)
    ap.add_argument("--guard", type=int, default=SkipOptions.guard,
                    help="candidate guard of the recorded tiles (4 or 0)")
    ap.add_argument("--workload", default="batch", choices=["batch", "interface"],
                    help="batch: the 256-beamlet workload's construction with other seeds; "
                         "interface: the bone-interface beam on the whole grid at seeded "
                         "beam offsets")
    ap.add_argument("--offsets", type=int, default=4,
                    help="interface: number of seeded beam offsets, U(-4, 4) mm")
    ap.add_argument("--out", type=Path, default=None)
    args = ap.parse_args(argv)
    if args.out is None:
        args.out = WEIGHTS_FILE if args.workload == "batch" else INTERFACE_WEIGHTS_FILE

    t0 = time.perf_counter()
    if args.workload == "batch":
        cfg = batch_workload_config()
        phantom, beamlets = training_workload(args.beamlets)
        handle, _ = prepare_batch(phantom, beamlets, cfg)
        num = _native.numerics_struct(cfg)
        x, y, near = record_samples(handle, num, horizon=args.horizon, guard=args.guard)
    else:
        x, y, near = record_interface_samples(interface_offsets(args.offsets), guard=args.guard,
                                              horizon=args.horizon)
    rec_s = time.perf_counter() - t0
    print(json.dumps({"recorded": len(y), "empty_fraction": float(y.mean()),
                      "near_fraction": float(near.mean()), "seconds": round(rec_s, 1)}),
          flush=True)
    rng = np.random.default_rng(args.seed)
    order = rng.permutation(len(y))
    n_test = len(y) // 10
    test, fit = order[:n_test], order[n_test:]
    t0 = time.perf_counter()
    model, loss = train(x[fit], y[fit], epochs=args.epochs, batch=args.batch, lr=args.lr,
                        seed=args.seed)
    report = {"train_loss": loss, "train_seconds": round(time.perf_counter() - t0, 1),
              "held_out": confusion(model, x[test], y[test], args.decision),
              "held_out_against_guard_rule": against_guard_rule(model, x[test], y[test],
                                                                near[test])}
    save_params(model.get_params(), args.out)
    report["weights"] = str(args.out)
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
