"""Spatial operator on the MI355X (drop-in for ``numerics/spatial.py:104-252``).

``spatial_operator`` returns -div F of the HLL fluxes of the WENO-Z face
states for a host state, computed by the fused stage kernel of
libm1march.so in its right-hand-side mode.
"""

from __future__ import annotations

import numpy as np

from .. import _native
from ..enums import SpatialScheme, WaveSpeeds


def spatial_operator(
    U,
    energy: float,
    grid,
    inflow,
    spatial_scheme: SpatialScheme,
    wave_speed_mode: WaveSpeeds,
    xp=np,
    active_domain: bool = False,
    active_relative_tolerance: float = 1.0e-14,
    boundary_state=None,
    *,
    material_map=None,
    config=None,
):
    """Negative divergence of the M1 spatial flux.

    The reference passes no material map here because its spatial operator
    does not depend on the material; the device context still needs one to be
    bound, so a homogeneous water map is used when none is given.
    ``active_domain`` crops are a host-side optimisation of the reference and
    are ignored: the full domain is evaluated.
    """
    from ..config import Numerics, Physics, SolverConfig
    from ..enums import Material
    from ..geometry import Geometry
    from ..physics.materials import MaterialMap
    from .energy import device_context

    if spatial_scheme is not SpatialScheme.WENO_Z:
        raise NotImplementedError(f"Spatial scheme {spatial_scheme} is not available on the MI355X.")
    U = np.ascontiguousarray(U.get() if hasattr(U, "get") else U, dtype=np.float64)
    shape = tuple(int(s) for s in grid.shape)
    if material_map is None:
        material_map = MaterialMap(ids=np.zeros(shape, dtype=np.int32), materials=(Material.WATER,))
    ctx = device_context(0 if config is None else config.device.device)
    ctx.set_problem(grid, material_map, inflow, inflow.beam)
    ctx.set_state(U)
    if config is None:
        # Only the wave-speed mode of the numerics is read by this operator.
        nx, ny, nz = shape
        config = SolverConfig(
            geometry=Geometry(size=(grid.dx * nx, grid.dy * ny, grid.dz * nz), resolution=shape),
            beam=inflow.beam,
            physics=Physics(),
            numerics=Numerics(wave_speeds=wave_speed_mode, energy_min=0.0),
        )
    num = _native.numerics_struct(config)
    num.wave_speeds = 0 if wave_speed_mode is WaveSpeeds.EXACT else 1
    if boundary_state is not None:
        boundary_state = np.asarray(boundary_state.get() if hasattr(boundary_state, "get") else boundary_state)
    rhs = ctx.spatial_operator(num, energy, boundary_state)
    return rhs if xp is np else xp.asarray(rhs)
