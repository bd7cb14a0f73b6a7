"""ctypes front end of the CPU oracle (liboracle_m1.so).

TEST INFRASTRUCTURE ONLY.  This module is the parity checker for the HIP
march and the timed CPU baseline of bench.py.  Only tests/, the smoke() entry
point and bench.py's cpu_baseline leg import it; the product never does.

``OracleProblem`` takes the same host objects the reference's ``energy_march``
receives (grid, material map, inflow, config) and runs the C restatement of
the reference march on them.  Material tables are read from the package's
data files as raw (energy, value) columns; the oracle computes its own PCHIP
slopes (``or_pchip_slopes``).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_m1.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class or_table(C.Structure):
    _fields_ = [("n", C.c_int), ("E", _dp), ("v", _dp), ("slope", _dp)]


class or_material(C.Structure):
    _fields_ = [("sp", or_table), ("tr", or_table), ("nu", or_table),
                ("density", C.c_double), ("radiation_length", C.c_double)]


class or_problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("ids", _ip), ("n_mat", C.c_int), ("mats", C.POINTER(or_material)),
                ("spatial", _dp), ("beam_energy", C.c_double),
                ("energy_spread", C.c_double), ("angular_spread", C.c_double)]


class or_config(C.Structure):
    _fields_ = [("energy_scheme", C.c_int), ("wave_speeds", C.c_int),
                ("scattering", C.c_int), ("nuclear", C.c_int),
                ("coefficient_interpolation", C.c_int),
                ("max_energy_step", C.c_double), ("cfl", C.c_double),
                ("energy_min", C.c_double), ("energy_window_sigma", C.c_double)]


def build() -> Path:
    """Compile liboracle_m1.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C {HERE}`."
            )
        L = C.CDLL(str(LIB_PATH))
        L.or_last_error.restype = C.c_char_p
        L.or_weno_z_left.restype = C.c_double
        L.or_weno_z_left.argtypes = [C.c_double] * 5
        L.or_linear_eval.restype = C.c_double
        L.or_linear_eval.argtypes = [C.POINTER(or_table), C.c_double]
        L.or_pchip_eval.argtypes = [C.POINTER(or_table), C.c_double, _dp]
        L.or_coefficient.argtypes = [C.POINTER(or_table), C.c_double, C.c_int, _dp]
        L.or_pchip_slopes.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_scattering_depth.argtypes = [C.POINTER(or_material), C.c_int, C.c_int,
                                          C.c_double, C.c_double, _dp]
        L.or_nuclear_depth.argtypes = [C.POINTER(or_material), C.c_int,
                                       C.c_double, C.c_double, _dp]
        L.or_project_realizable.argtypes = [_dp, C.c_long, C.c_double]
        L.or_eigenvalues.argtypes = [_dp, C.c_long, C.c_int, _dp, _dp]
        L.or_hll_flux.argtypes = [_dp, _dp, C.c_long, C.c_int, C.c_int, _dp]
        L.or_fill_ghost_cells.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_reconstruct.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        P = C.POINTER(or_problem)
        Cf = C.POINTER(or_config)
        L.or_spatial_operator.argtypes = [P, C.c_int, _dp, C.c_double, _dp, _dp]
        L.or_inflow_state.argtypes = [P, C.c_double, _dp]
        L.or_compute_energy_step.argtypes = [P, Cf, _dp, C.c_double, _dp]
        for name in ("or_forward_euler", "or_ssprk3", "or_if_ssprk3", "or_strang_step"):
            getattr(L, name).argtypes = [P, Cf, _dp, C.c_double, C.c_double, _dp]
        L.or_energy_march.argtypes = [P, Cf, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int),
                                      _dp, C.c_int, _dp, _ip, C.c_int]
        L.or_compute_energy_step_vacuum.argtypes = [P, Cf, _dp, C.c_double, C.c_double, _dp]
        L.or_energy_march_moving.argtypes = [P, Cf, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                             _dp, _dp, C.c_int, C.POINTER(C.c_int), _dp,
                                             C.c_int, _dp, _ip, C.c_int, _ip]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_rk3_trace_combination.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().or_last_error().decode())


def set_threads(n: int):
    lib().or_set_threads(int(n))


def make_table(E, v, with_slopes: bool):
    """(or_table, keep-alive list) with oracle-computed PCHIP slopes."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    keep = [E, v]
    slope = None
    if with_slopes:
        slope = np.empty_like(v)
        _check(lib().or_pchip_slopes(len(E), _ptr(E), _ptr(v), _ptr(slope)))
        keep.append(slope)
    return or_table(len(E), _ptr(E), _ptr(v), _ptr(slope)), keep


_SCHEME = {"FORWARD_EULER": 0, "SSPRK3": 1, "IF_SSPRK3": 2}
_WAVE = {"EXACT": 0, "UNIT": 1}
_SCAT = {"OFF": 0, "ROSSI": 1, "WENTZEL": 2}
_INTERP = {"LINEAR": 0, "PCHIP": 1}


def make_config(config) -> or_config:
    n = config.numerics
    ph = config.physics
    if n.energy_scheme.name not in _SCHEME:
        raise NotImplementedError(f"Energy scheme {n.energy_scheme} has not been migrated yet.")
    return or_config(_SCHEME[n.energy_scheme.name], _WAVE[n.wave_speeds.name],
                     _SCAT[ph.scattering.name], int(bool(ph.nuclear)),
                     _INTERP[ph.coefficient_interpolation.name],
                     float(n.max_energy_step), float(n.cfl), float(n.energy_min),
                     float(n.energy_window_sigma))


def material_struct(material):
    """or_material for a proton_solver Material, from the raw data tables."""
    from proton_solver.physics.materials import density, radiation_length
    from proton_solver.physics.stopping_power import load_stopping_power_table
    from proton_solver.physics.scattering import load_transport_table
    from proton_solver.physics.nuclear import load_nuclear_table

    sp, k1 = make_table(*load_stopping_power_table(material), True)
    tr, k2 = make_table(*load_transport_table(material), True)
    nu, k3 = make_table(*load_nuclear_table(material), True)
    m = or_material(sp, tr, nu, density(material), radiation_length(material))
    return m, k1 + k2 + k3


class OracleProblem:
    """One problem (grid, materials, inflow, config) bound to the oracle."""

    def __init__(self, grid, material_map, inflow, config):
        self.grid = grid
        self.config = config
        self.shape = tuple(int(s) for s in grid.shape)
        nx, ny, nz = self.shape
        self.ids = np.ascontiguousarray(np.asarray(material_map.ids), dtype=np.int32)
        self.spatial = np.ascontiguousarray(np.asarray(inflow.spatial), dtype=np.float64)
        self._keep = []
        self.materials = list(material_map.materials)
        mats = (or_material * len(material_map.materials))()
        for i, m in enumerate(material_map.materials):
            mats[i], keep = material_struct(m)
            self._keep += keep
        self._mats = mats
        b = config.beam
        self.prob = or_problem(nx, ny, nz, float(grid.dx), float(grid.dy), float(grid.dz),
                               _ptr(self.ids), len(material_map.materials), mats,
                               _ptr(self.spatial), float(b.energy), float(b.energy_spread),
                               float(b.angular_spread))
        self.cfg = make_config(config)

    @classmethod
    def from_config(cls, config):
        from proton_solver.numerics.boundary import BeamInflow
        from proton_solver.numerics.grid import Grid
        from proton_solver.physics.materials import MaterialMap

        grid = Grid.from_geometry(config.geometry, np)
        mm = MaterialMap.from_geometry(config.geometry, grid, np)
        inflow = BeamInflow.from_beam(config.beam, grid, np)
        return cls(grid, mm, inflow, config)

    @property
    def n(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    def zeros(self):
        return np.zeros((4, *self.shape))

    def inflow_state(self, energy):
        out = np.empty((4, self.shape[0], self.shape[1]))
        _check(lib().or_inflow_state(C.byref(self.prob), energy, _ptr(out)))
        return out

    def spatial_operator(self, U, energy, boundary_state=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        b = None if boundary_state is None else np.ascontiguousarray(boundary_state, np.float64)
        rhs = np.empty_like(U)
        _check(lib().or_spatial_operator(C.byref(self.prob), self.cfg.wave_speeds,
                                         _ptr(U), energy, _ptr(b), _ptr(rhs)))
        return rhs

    def compute_energy_step(self, U, energy):
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = C.c_double()
        _check(lib().or_compute_energy_step(C.byref(self.prob), C.byref(self.cfg),
                                            _ptr(U), energy, C.byref(out)))
        return out.value

    def compute_energy_step_vacuum(self, U, energy, relative_tolerance):
        """compute_energy_step with vacuum_relative_tolerance (energy.py:459-660)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = C.c_double()
        _check(lib().or_compute_energy_step_vacuum(C.byref(self.prob), C.byref(self.cfg),
                                                   _ptr(U), energy, float(relative_tolerance),
                                                   C.byref(out)))
        return out.value

    def march_moving(self, U=None, snapshot_energies=(), max_steps=0, guard=None,
                     truncation_sigma=None, vacuum_aware=None):
        """_moving_active_energy_march: returns (U, dose [Gy], energies, snapshots,
        boxes), boxes[s] = (i0, i1, j0, j1, k0, k1) of step s.  guard, sigma and
        the vacuum-aware CFL default to the config's numerics."""
        num = self.config.numerics
        guard = num.active_guard_cells if guard is None else guard
        sigma = num.active_truncation_sigma if truncation_sigma is None else truncation_sigma
        va = num.vacuum_aware_cfl if vacuum_aware is None else vacuum_aware
        U0 = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        U_out = np.empty((4, *self.shape))
        dose = np.empty(self.shape)
        cap = 1 << 16
        energies = np.empty(cap)
        boxes = np.zeros((cap, 6), dtype=np.int32)
        ne = C.c_int(0)
        snaps = np.ascontiguousarray([float(s) for s in snapshot_energies], dtype=np.float64)
        ns = len(snaps)
        snap_out = np.empty((max(ns, 1), 4, *self.shape)) if ns else None
        taken = np.zeros(max(ns, 1), dtype=np.int32)
        _check(lib().or_energy_march_moving(
            C.byref(self.prob), C.byref(self.cfg), int(guard), float(sigma), int(bool(va)),
            _ptr(U0), _ptr(U_out), _ptr(dose), _ptr(energies), cap, C.byref(ne),
            _ptr(snaps) if ns else None, ns, _ptr(snap_out), _ptr(taken), int(max_steps),
            _ptr(boxes)))
        snapshots = {}
        for i in range(ns):
            if taken[i]:
                snapshots[float(snaps[i])] = snap_out[i].copy()
        k = min(ne.value, cap)
        return U_out, dose, energies[:k].copy(), snapshots, boxes[: max(k - 1, 0)].copy()

    def step(self, U, energy, dE, kind="if_ssprk3"):
        """One integrator step; returns (U_new, energy_deposition)."""
        U = np.array(U, dtype=np.float64, order="C", copy=True)
        edep = np.empty(self.shape)
        fn = getattr(lib(), "or_" + kind)
        _check(fn(C.byref(self.prob), C.byref(self.cfg), _ptr(U), energy, dE, _ptr(edep)))
        return U, edep

    def march(self, U=None, snapshot_energies=(), max_steps=0):
        """energy_march: returns (U, dose [Gy], energies, snapshots)."""
        U = self.zeros() if U is None else np.array(U, dtype=np.float64, order="C", copy=True)
        dose = np.empty(self.shape)
        cap = 1 << 20
        energies = np.empty(cap)
        ne = C.c_int(0)
        snaps = np.ascontiguousarray([float(s) for s in snapshot_energies], dtype=np.float64)
        ns = len(snaps)
        snap_out = np.empty((max(ns, 1), 4, *self.shape)) if ns else None
        taken = np.zeros(max(ns, 1), dtype=np.int32)
        _check(lib().or_energy_march(C.byref(self.prob), C.byref(self.cfg), _ptr(U), _ptr(dose),
                                     _ptr(energies), cap, C.byref(ne),
                                     _ptr(snaps) if ns else None, ns,
                                     _ptr(snap_out), _ptr(taken), int(max_steps)))
        snapshots = {}
        for i in range(ns):
            if taken[i]:
                snapshots[float(snaps[i])] = snap_out[i].copy()
        return U, dose, energies[: min(ne.value, cap)].copy(), snapshots


def weno_z_left(v0, v1, v2, v3, v4) -> float:
    return lib().or_weno_z_left(v0, v1, v2, v3, v4)


def project(U, eps=1.0e-10):
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    n = U.size // 4
    lib().or_project_realizable(_ptr(U), n, eps)
    return U


def eigenvalues(U, axis):
    U = np.ascontiguousarray(U, dtype=np.float64)
    n = U.size // 4
    lo = np.empty(U.shape[1:])
    hi = np.empty(U.shape[1:])
    lib().or_eigenvalues(_ptr(U), n, axis, _ptr(lo), _ptr(hi))
    return lo, hi


def hll_flux(UL, UR, axis, wave_mode=0):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    F = np.empty_like(UL)
    lib().or_hll_flux(_ptr(UL), _ptr(UR), UL.size // 4, axis, wave_mode, _ptr(F))
    return F


def fill_ghost_cells(U, U_in):
    U = np.ascontiguousarray(U, dtype=np.float64)
    U_in = np.ascontiguousarray(U_in, dtype=np.float64)
    _, nx, ny, nz = U.shape
    out = np.empty((4, nx + 6, ny + 6, nz + 6))
    _check(lib().or_fill_ghost_cells(_ptr(U), nx, ny, nz, _ptr(U_in), _ptr(out)))
    return out


def reconstruct(U_ext, axis):
    U_ext = np.ascontiguousarray(U_ext, dtype=np.float64)
    _, ex, ey, ez = U_ext.shape
    nx, ny, nz = ex - 6, ey - 6, ez - 6
    shape = [4, nx, ny, nz]
    shape[axis + 1] += 1
    L = np.empty(shape)
    R = np.empty(shape)
    lib().or_reconstruct(_ptr(U_ext), nx, ny, nz, axis, _ptr(L), _ptr(R))
    return L, R


def rk3_trace_combination(B0, B13, B23, B1):
    arrs = [np.ascontiguousarray(b, dtype=np.float64) for b in (B0, B13, B23, B1)]
    o1 = np.empty_like(arrs[0])
    o2 = np.empty_like(arrs[0])
    lib().or_rk3_trace_combination(arrs[0].size, *[_ptr(a) for a in arrs], _ptr(o1), _ptr(o2))
    return o1, o2


def material_from_arrays(sp, tr, nu, density, radiation_length):
    """or_material from (E, v) pairs; slopes computed by the oracle."""
    keep = []
    tabs = []
    for E, v in (sp, tr, nu):
        t, k = make_table(E, v, True)
        tabs.append(t)
        keep += k
    return or_material(*tabs, float(density), float(radiation_length)), keep


def scattering_depth(mat, model, energy, dE, interp="LINEAR"):
    out = C.c_double()
    _check(lib().or_scattering_depth(C.byref(mat), _SCAT[model], _INTERP[interp], energy, dE,
                                     C.byref(out)))
    return out.value


def nuclear_depth(mat, energy, dE, interp="LINEAR"):
    out = C.c_double()
    _check(lib().or_nuclear_depth(C.byref(mat), _INTERP[interp], energy, dE, C.byref(out)))
    return out.value


def table_eval(E, v, energy, interp="PCHIP"):
    t, keep = make_table(E, v, True)
    out = C.c_double()
    _check(lib().or_coefficient(C.byref(t), float(energy), _INTERP[interp], C.byref(out)))
    return out.value


def pchip_slopes(E, v):
    E = np.ascontiguousarray(E, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    s = np.empty_like(v)
    _check(lib().or_pchip_slopes(len(E), _ptr(E), _ptr(v), _ptr(s)))
    return s


def threads_env() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
