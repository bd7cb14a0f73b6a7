"""Solver front end (mirrors ``proton_solver/solver.py:29-217``).

The deterministic march runs on the MI355X.  The reference's comparison
references (analytical CSDA, BFP Monte Carlo, TOPAS) are outside this package;
a request for one is reported with a warning and its dose stays None.
"""

from __future__ import annotations

import time
import warnings

from .backend import resolve_backend
from .config import SolverConfig
from .enums import Reference
from .numerics.boundary import BeamInflow
from .numerics.energy import energy_march
from .numerics.grid import Grid
from .physics.materials import MaterialMap
from .result import Result


class Solver:
    def __init__(self, config: SolverConfig):
        self.config = config
        self.backend = resolve_backend(config.backend)

    def solve(self, snapshot_energies: tuple[float, ...] = ()) -> Result:
        cfg = self.config
        xp = self.backend.xp
        grid = Grid.from_geometry(cfg.geometry, xp)
        material_map = MaterialMap.from_geometry(cfg.geometry, grid, xp)
        inflow = BeamInflow.from_beam(cfg.beam, grid, xp)

        phi = J = dose = energies = None
        snapshots = {}
        runtime = None
        if cfg.solve_deterministic:
            # The moving active domain starts from U = None (solver.py:90-100).
            U = None if cfg.numerics.moving_active_domain else xp.zeros((4, *grid.shape),
                                                                        dtype=float)
            start = time.perf_counter()
            U, dose, energies, snapshots = energy_march(
                U, material_map, grid, inflow, cfg, xp, snapshot_energies=snapshot_energies
            )
            runtime = time.perf_counter() - start
            if U is not None and cfg.output.store_final_state:
                phi, J = U[0], U[1:4]

        for ref in (cfg.references.reference, *cfg.references.additional_references):
            if ref is not Reference.NONE:
                warnings.warn(
                    f"Reference {ref.name} is not computed by the MI355X march package.",
                    stacklevel=2,
                )
        result = Result(
            phi=phi, J=J, dose=dose, grid=grid, energies=energies, snapshots=snapshots,
            runtime=runtime, backend=self.backend.backend,
        )
        if cfg.output.save_results:
            from .persistence import save_result

            result.run_directory = save_result(result, cfg)
        return result
