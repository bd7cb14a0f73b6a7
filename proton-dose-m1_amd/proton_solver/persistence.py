"""Result persistence (mirrors ``proton_solver/persistence.py:13-369``).

A saved run is a directory ``<timestamp>__<nx>x<ny>x<nz>__<scheme>__<config
hash[:8]>__<git state hash[:8] | nogit>`` under ``output.results_directory``
holding

* ``config.json``   the canonical configuration (``config_dict``),
* ``metadata.json`` creation time, configuration hash, git provenance,
  runtimes, speed-ups and backend,
* ``results.npz``   x, y, z and, when present, phi, J, dose, energies and
  the reference doses (``numpy.savez_compressed``; loads without pickling).

The configuration hash is the SHA-256 of ``config_dict`` dumped as JSON with
sorted keys and no whitespace (persistence.py:73-85), so runs of the same
configuration share it with the reference's runs.  ``config_dict`` drops the
output settings and ``geometry.view_geometry`` as the reference does, and
also the MI355X device options when they are the defaults (FP64 on device
0, which computes the reference's march); other device options change the
arithmetic and stay in the hash.
"""

from __future__ import annotations

import hashlib
import json
import subprocess
from dataclasses import fields, is_dataclass
from datetime import datetime
from enum import Enum
from pathlib import Path
from typing import Any

import numpy as np


def _to_serializable(value: Any):
    """JSON-compatible form of a configuration value (persistence.py:13-53)."""
    if isinstance(value, Enum):
        return value.name
    if isinstance(value, np.ndarray):
        # volumes (CT) by content digest, not element by element
        return {
            "shape": [int(n) for n in value.shape],
            "dtype": str(value.dtype),
            "sha256": hashlib.sha256(np.ascontiguousarray(value).tobytes()).hexdigest(),
        }
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, Path):
        return str(value)
    if is_dataclass(value) and not isinstance(value, type):
        return {f.name: _to_serializable(getattr(value, f.name)) for f in fields(value)}
    if isinstance(value, dict):
        return {str(k): _to_serializable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_to_serializable(v) for v in value]
    return value


def config_dict(config) -> dict:
    """Canonical simulation configuration (persistence.py:56-70)."""
    from .config import DeviceOptions

    data = _to_serializable(config)
    data.pop("output", None)
    if "geometry" in data:
        data["geometry"].pop("view_geometry", None)
    device = getattr(config, "device", None)
    if device is None or device == DeviceOptions(device=device.device):
        data.pop("device", None)
    return data


def hash_config_dict(data: dict) -> str:
    """SHA-256 of a configuration dictionary in its canonical JSON form."""
    payload = json.dumps(data, sort_keys=True, separators=(",", ":")).encode("utf-8")
    return hashlib.sha256(payload).hexdigest()


def config_hash(config) -> str:
    """Stable SHA-256 of the simulation configuration (persistence.py:73-85)."""
    return hash_config_dict(config_dict(config))


def _git(*args, cwd=None):
    return subprocess.run(["git", *args], capture_output=True, check=True, cwd=cwd).stdout


def git_provenance(cwd=None) -> dict:
    """Commit, dirty flag and a digest of the working tree
    (persistence.py:88-180): commit id, ``git diff HEAD --binary`` and the
    untracked files (names and contents) hashed together."""
    try:
        commit = _git("rev-parse", "HEAD", cwd=cwd).decode().strip()
        top = Path(_git("rev-parse", "--show-toplevel", cwd=cwd).decode().strip())
        diff = _git("diff", "HEAD", "--binary", cwd=cwd)
        untracked = _git("ls-files", "--others", "--exclude-standard",
                         cwd=cwd).decode().splitlines()
        h = hashlib.sha256()
        h.update(commit.encode("utf-8"))
        h.update(diff)
        for rel in sorted(untracked):
            h.update(rel.encode("utf-8"))
            path = top / rel
            if path.is_file():
                h.update(path.read_bytes())
        return {"git_commit": commit, "git_dirty": bool(diff or untracked),
                "git_state_hash": h.hexdigest()}
    except (subprocess.CalledProcessError, FileNotFoundError):
        return {"git_commit": None, "git_dirty": None, "git_state_hash": None}


def _host(a):
    if a is None:
        return None
    return a.get() if hasattr(a, "get") else np.asarray(a)


def _speedup(ref_runtime, runtime):
    if runtime is None or ref_runtime is None or runtime <= 0.0 or ref_runtime <= 0.0:
        return None
    return float(ref_runtime / runtime)


def save_result(result, config, git_cwd=None) -> Path:
    """Write config.json, metadata.json and results.npz of one run
    (persistence.py:195-369); returns the run directory."""
    root = Path(config.output.results_directory)
    root.mkdir(parents=True, exist_ok=True)
    digest = config_hash(config)
    git = git_provenance(git_cwd)
    git_tag = git["git_state_hash"][:8] if git["git_state_hash"] is not None else "nogit"
    now = datetime.now()
    nx, ny, nz = config.geometry.resolution
    scheme = config.numerics.spatial_scheme.name.lower()
    run = root / f"{now.strftime('%Y%m%d_%H%M%S')}__{nx}x{ny}x{nz}__{scheme}__{digest[:8]}__{git_tag}"
    run.mkdir()

    with (run / "config.json").open("w") as f:
        json.dump(config_dict(config), f, indent=2, sort_keys=True)

    label = result.reference_label
    metadata = {
        "created_at": now.astimezone().isoformat(),
        "config_hash": digest,
        **git,
        "runtime_seconds": result.runtime,
        "bfp_mc_runtime_seconds": result.bfp_mc_runtime,
        "topas_runtime_seconds": result.topas_runtime,
        "speedup_vs_bfp_mc_x": ("-" if label != "BFP Monte Carlo"
                                else _speedup(result.bfp_mc_runtime, result.runtime)),
        "speedup_vs_topas_x": ("-" if label != "TOPAS"
                               else _speedup(result.topas_runtime, result.runtime)),
        "backend": result.backend.name if result.backend is not None else None,
    }
    with (run / "metadata.json").open("w") as f:
        json.dump(metadata, f, indent=2)

    arrays = {"x": _host(result.x), "y": _host(result.y), "z": _host(result.z)}
    for name in ("phi", "J", "dose", "energies", "analytical_dose", "topas_dose",
                 "bfp_mc_dose"):
        value = getattr(result, name)
        if value is not None:
            arrays[name] = _host(value)
    np.savez_compressed(run / "results.npz", **arrays)
    return run
