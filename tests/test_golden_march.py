"""The oracle's whole march against the reference's published FP64 results.

tests/golden/thesis_golden.json holds the scalar dose metrics that the
reference's own Numba FP64 march produced for the coarse-grid thesis runs
(thesis_results/robustness/coarse_grid_{2mm,1mm}/*/*/weno_z/metrics.json,
collected by tests/golden/make_thesis_fixtures.py).  CPU only.

2 mm runs: every energy step is capped by max_energy_step, so the step
sequence is independent of rounding and the oracle must reproduce maximum
dose to 5e-9 (observed <= 1.4e-9: the reference runs WENO-Z and HLL under
Numba fastmath), R80 and DFW to 1e-6 mm, and the Bragg-peak depth exactly.

1 mm runs: the CFL bound is active and its maximum is attained in cells
whose speeds are roundoff noise, so any change of rounding changes the
energy grid (the oracle built with FMA contraction moves from 626 to 628
steps).  These runs are compared to 2e-3 and only with M1_SLOW=1.
"""

import os

import numpy as np
import pytest

from helpers import config_from_fixture, thesis_cases
from proton_solver import metrics

CASES_2MM = [c for c in thesis_cases() if "coarse_grid_2mm" in c["source"]]
CASES_1MM = [c for c in thesis_cases() if "coarse_grid_1mm" in c["source"]]


def _run(O, case):
    cfg = config_from_fixture(case["config"])
    op = O.OracleProblem.from_config(cfg)
    U, dose, E, _ = op.march()
    return op.grid, U, dose, E


def _ids(cases):
    return [c["source"].split("robustness/")[1].replace("/weno_z", "") for c in cases]


@pytest.mark.parametrize("case", CASES_2MM, ids=_ids(CASES_2MM))
def test_oracle_reproduces_published_fp64_march_2mm(oracle_lib, case):
    grid, U, dose, E = _run(oracle_lib, case)
    assert len(E) - 1 == 620
    assert abs(dose.max() / case["maximum_dose_Gy"] - 1) < 5e-9
    r80, dfw = metrics.range_metrics(dose, grid.z, case["peak_search_from_mm"] / 1000)
    assert abs(r80 * 1000 - case["r80_mm"]) < 1e-6
    assert abs(dfw * 1000 - case["dfw_mm"]) < 1e-6
    assert metrics.peak_depth(dose, grid.x, grid.y, grid.z) * 1000 == pytest.approx(
        case["central_axis_peak_depth_mm"], abs=1e-9)
    assert np.all(np.isfinite(U)) and np.all(dose >= 0)


@pytest.mark.skipif(os.environ.get("M1_SLOW") != "1", reason="set M1_SLOW=1 (about 3 min)")
@pytest.mark.parametrize("case", CASES_1MM, ids=_ids(CASES_1MM))
def test_oracle_close_to_published_fp64_march_1mm(oracle_lib, case):
    grid, U, dose, E = _run(oracle_lib, case)
    assert abs(dose.max() / case["maximum_dose_Gy"] - 1) < 2e-3
    assert metrics.peak_depth(dose, grid.x, grid.y, grid.z) * 1000 == pytest.approx(
        case["central_axis_peak_depth_mm"], abs=1e-9)
