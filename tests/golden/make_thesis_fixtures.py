"""Collect the reference's published FP64 march results into a small fixture.

The reference ships, for every thesis run, the run configuration
(config.json) and the scalar dose metrics of its own Numba FP64 march
(metrics.json) under thesis_results/.  This script copies those numbers -- no
reference code is imported or executed -- into thesis_golden.json, which the
tests use to pin the oracle.  Run it in a checkout that has the reference at
REFERENCE_ROOT (default /root/reference).
"""

import json
import os
from pathlib import Path

ROOT = Path(os.environ.get("REFERENCE_ROOT", "/root/reference"))
OUT = Path(__file__).resolve().parent / "thesis_golden.json"

STUDIES = [
    "thesis_results/robustness/coarse_grid_2mm",
    "thesis_results/robustness/coarse_grid_1mm",
    "thesis_results/robustness/coarse_grid",            # 50 x 50 x 200 cells
    # 100 x 100 x 400 cells: static and moving active domain
    "thesis_results/reduced_physics",
    "thesis_results/scattering_nuclear",
    "thesis_results/robustness/moving_active_domain",
]
MOVING_KEYS = ("active_guard_cells", "active_truncation_sigma", "vacuum_aware_cfl")


def main():
    cases = []
    for study in STUDIES:
        base = ROOT / study
        for metrics_path in sorted(base.glob("**/weno_z/metrics.json")):
            run = metrics_path.parent
            cfg = json.loads((run / "config.json").read_text())
            met = json.loads(metrics_path.read_text())
            meta = json.loads((run / "metadata.json").read_text())
            cases.append({
                "source": str(run.relative_to(ROOT)),
                "config": {
                    "beam": cfg["beam"],
                    "geometry": {
                        k: cfg["geometry"][k]
                        for k in ("size", "resolution", "background_material", "objects")
                    },
                    "physics": cfg["physics"],
                    "numerics": {
                        k: cfg["numerics"][k]
                        for k in ("energy_scheme", "spatial_scheme", "wave_speeds", "cfl",
                                  "energy_min", "energy_window_sigma", "max_energy_step",
                                  "moving_active_domain", *MOVING_KEYS)
                        if k in cfg["numerics"] and (k not in MOVING_KEYS
                                                     or cfg["numerics"]["moving_active_domain"])
                    },
                },
                "maximum_dose_Gy": met["maximum_dose"]["deterministic_Gy"],
                "r80_mm": met["range"]["deterministic_r80_mm"],
                "dfw_mm": met["range"]["deterministic_dfw_mm"],
                "peak_search_from_mm": met["range"].get("peak_search_from_mm"),
                "central_axis_peak_depth_mm": met["central_axis_bragg_peak"]["deterministic_depth_mm"],
                "runtime_seconds": meta.get("runtime_seconds"),
                # the stored configuration as written and its hash (persistence)
                "config_json": cfg,
                "config_hash": meta.get("config_hash"),
            })
    OUT.write_text(json.dumps({"cases": cases}, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
