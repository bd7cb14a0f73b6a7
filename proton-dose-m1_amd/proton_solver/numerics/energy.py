"""Energy march on the MI355X (drop-in for ``numerics/energy.py``).

``energy_march`` keeps the reference signature and return contract
(energy.py:1497-1644): it takes the host state, material map, grid, inflow and
config, marches from ``E_beam + energy_window_sigma * sigma_E`` down to
``energy_min`` and returns ``(U, dose, energies, snapshots)``.  The march runs
entirely in libm1march.so; there is no host fallback.

The single-step functions (``forward_euler``, ``ssprk3``, ``if_ssprk3``,
``strang_energy_step``) and ``compute_energy_step`` mirror the reference
functions of the same names and also run on the device.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..enums import Scattering, SpatialScheme

MEV_TO_JOULE = 1.602176634e-13

_contexts: dict[int, _native.Context] = {}


def device_context(device: int = 0) -> _native.Context:
    """Process-wide libm1march context of one device."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = _native.Context(device)
        _contexts[device] = ctx
    return ctx


def _check_supported(config):
    if config.numerics.spatial_scheme is not SpatialScheme.WENO_Z:
        raise NotImplementedError(
            f"Spatial scheme {config.numerics.spatial_scheme} is not available on the "
            "MI355X march (WENO_Z only)."
        )


def _bind(U, material_map, grid, inflow, config):
    _check_supported(config)
    ctx = device_context(getattr(getattr(config, "device", None), "device", 0))
    ctx.set_problem(grid, material_map, inflow, config.beam)
    ctx.set_state(None if U is None else np.asarray(U))
    return ctx


def _host(a):
    return a.get() if hasattr(a, "get") else np.asarray(a)


def energy_march(
    U: Any,
    material_map,
    grid,
    inflow,
    config,
    xp=np,
    snapshot_energies: tuple[float, ...] = (),
):
    """March the M1 system from high to low kinetic energy (energy.py:1497-1644)."""
    if config.numerics.moving_active_domain:
        return _moving_active_energy_march(U, material_map, grid, inflow, config, xp,
                                           snapshot_energies)
    if U is None:
        raise ValueError(
            "A full-grid initial state is required when moving_active_domain is disabled."
        )
    ctx = _bind(_host(U), material_map, grid, inflow, config)
    num = _native.numerics_struct(config)
    energies, snapshots = ctx.march(num, snapshot_energies)
    U_out = ctx.get_state()
    dose = ctx.get_dose()
    if xp is not np:
        U_out, dose = xp.asarray(U_out), xp.asarray(dose)
        snapshots = {k: xp.asarray(v) for k, v in snapshots.items()}
    return U_out, dose, energies, snapshots


def _moving_active_energy_march(U, material_map, grid, inflow, config, xp=np,
                                snapshot_energies=()):
    """March on a box that follows the protons (energy.py:1262-1494).

    The state lives on the bounding box of the cells with
    phi > exp(-active_truncation_sigma^2 / 2) * max phi, widened by
    active_guard_cells and re-found before every step; the dose is kept on the
    whole grid.  U may be None (zero start, as Solver passes it) or a
    full-grid state, which is cut to the initial box.  The final state is the
    box state embedded in the grid, or None when output.store_final_state is
    False.  The march runs in libm1march.so (m1_march_moving).
    """
    num_cfg = config.numerics
    ctx = _bind(None if U is None else _host(U), material_map, grid, inflow, config)
    num = _native.numerics_struct(config)
    energies, snapshots, _ = ctx.march_moving(num, num_cfg.active_guard_cells,
                                              num_cfg.active_truncation_sigma,
                                              snapshot_energies)
    dose = ctx.get_dose()
    store = getattr(getattr(config, "output", None), "store_final_state", True)
    U_out = ctx.get_state() if store else None
    if xp is not np:
        dose = xp.asarray(dose)
        U_out = None if U_out is None else xp.asarray(U_out)
        snapshots = {k: xp.asarray(v) for k, v in snapshots.items()}
    return U_out, dose, energies, snapshots


def compute_energy_step(U, energy, material_map, grid, inflow, config, xp=np, *,
                        vacuum_relative_tolerance=None):
    """Stable energy step from the spatial CFL condition (energy.py:539-660);
    with vacuum_relative_tolerance the vacuum-aware directional masks of
    energy.py:459-537 apply."""
    ctx = _bind(_host(U), material_map, grid, inflow, config)
    num = _native.numerics_struct(config)
    if vacuum_relative_tolerance is not None:
        return ctx.compute_energy_step_vacuum(num, energy, vacuum_relative_tolerance)
    return ctx.compute_energy_step(num, energy)


def _one_step(kind, U, energy, dE, material_map, grid, inflow, config, xp):
    ctx = _bind(_host(U), material_map, grid, inflow, config)
    edep = ctx.step(_native.numerics_struct(config), kind, energy, dE)
    U_new = ctx.get_state()
    if xp is not np:
        return xp.asarray(U_new), xp.asarray(edep)
    return U_new, edep


def forward_euler(U, energy, dE, material_map, grid, inflow, config, xp=np):
    """One forward-Euler step (energy.py:663-722)."""
    return _one_step(_native.STEP_FORWARD_EULER, U, energy, dE, material_map, grid, inflow,
                     config, xp)


def ssprk3(U, energy, dE, material_map, grid, inflow, config, xp=np):
    """One SSPRK3 step (energy.py:725-873)."""
    return _one_step(_native.STEP_SSPRK3, U, energy, dE, material_map, grid, inflow, config, xp)


def if_ssprk3(U, energy, dE, material_map, grid, inflow, config, xp=np):
    """One integrating-factor SSPRK3 step (energy.py:878-1130)."""
    return _one_step(_native.STEP_IF_SSPRK3, U, energy, dE, material_map, grid, inflow,
                     config, xp)


def strang_energy_step(U, energy, dE, material_map, grid, inflow, config, xp=np):
    """Strang-split source / transport step (energy.py:1194-1260)."""
    return _one_step(_native.STEP_STRANG, U, energy, dE, material_map, grid, inflow, config, xp)


def advance_energy_step(U, energy, dE, material_map, grid, inflow, config, xp=np):
    """Configured transport integrator (energy.py:1132-1165)."""
    from ..enums import EnergyScheme

    scheme = config.numerics.energy_scheme
    if scheme is EnergyScheme.FORWARD_EULER:
        return forward_euler(U, energy, dE, material_map, grid, inflow, config, xp)
    if scheme in (EnergyScheme.SSPRK3, EnergyScheme.IF_SSPRK3):
        return ssprk3(U, energy, dE, material_map, grid, inflow, config, xp)
    raise NotImplementedError(f"Energy scheme {scheme} has not been migrated yet.")


def uses_sources(config) -> bool:
    return config.physics.scattering is not Scattering.OFF or bool(config.physics.nuclear)
