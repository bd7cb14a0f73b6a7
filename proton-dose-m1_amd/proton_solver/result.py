"""Solver result (mirrors ``proton_solver/result.py:10-152``, deterministic part)."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

from .enums import Backend


@dataclass
class Result:
    phi: Any
    J: Any
    dose: Any

    grid: Any
    energies: Any
    snapshots: dict = field(default_factory=dict)

    analytical_dose: Any = None
    topas_dose: Any = None
    bfp_mc_dose: Any = None

    runtime: float | None = None
    topas_runtime: float | None = None
    bfp_mc_runtime: float | None = None
    backend: Backend | None = None
    run_directory: Any = None

    @property
    def x(self):
        return self.grid.x

    @property
    def y(self):
        return self.grid.y

    @property
    def z(self):
        return self.grid.z

    @property
    def primary_dose(self):
        if self.dose is not None:
            return self.dose
        for d in (self.bfp_mc_dose, self.topas_dose, self.analytical_dose):
            if d is not None:
                return d
        return None

    @property
    def reference_dose(self):
        for d in (self.bfp_mc_dose, self.topas_dose, self.analytical_dose):
            if d is not None:
                return d
        return None

    @property
    def reference_label(self):
        """Name of the selected reference (result.py:119-140)."""
        if (self.dose is None and self.analytical_dose is not None
                and (self.bfp_mc_dose is not None or self.topas_dose is not None)):
            return "Analytical"
        if self.bfp_mc_dose is not None:
            return "BFP Monte Carlo"
        if self.topas_dose is not None:
            return "TOPAS"
        if self.analytical_dose is not None:
            return "Analytical"
        return None

    @property
    def maximum_dose(self):
        import numpy as np

        return None if self.dose is None else float(np.max(self.dose))
