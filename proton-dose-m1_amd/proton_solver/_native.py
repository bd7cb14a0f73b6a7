"""ctypes binding of libm1march.so (declarations: include/m1march.h).

The library is the only compute path of this package: there is no CPU
fallback.  Loading fails loudly when the library is missing or was built for
another ABI version.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .enums import CoefficientInterpolation, EnergyScheme, Precision, Scattering, WaveSpeeds

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libm1march.so"
ABI_VERSION = 5

M1_OK = 0
M1_ERR_VALUE = -1
M1_ERR_NOT_IMPLEMENTED = -2
M1_ERR_HIP = -3
M1_ERR_NONFINITE = -4
M1_ERR_STATE = -5

STEP_FORWARD_EULER = 0
STEP_SSPRK3 = 1
STEP_IF_SSPRK3 = 2
STEP_STRANG = 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class m1_table(C.Structure):
    _fields_ = [("n", C.c_int32), ("energy", _dp), ("value", _dp), ("slope", _dp)]


class m1_material(C.Structure):
    _fields_ = [("stopping_power", m1_table), ("transport", m1_table), ("removal", m1_table),
                ("density", C.c_double), ("radiation_length", C.c_double)]


class m1_beam(C.Structure):
    _fields_ = [("energy", C.c_double), ("energy_spread", C.c_double),
                ("angular_spread", C.c_double)]


class m1_numerics(C.Structure):
    _fields_ = [("energy_scheme", C.c_int32), ("wave_speeds", C.c_int32),
                ("scattering", C.c_int32), ("nuclear", C.c_int32),
                ("coefficient_interpolation", C.c_int32), ("precision", C.c_int32),
                ("max_energy_step", C.c_double), ("cfl", C.c_double),
                ("energy_min", C.c_double), ("energy_window_sigma", C.c_double),
                ("tile_threshold", C.c_double), ("vacuum_aware_cfl", C.c_int32),
                ("batch_cfl_box_only", C.c_int32)]


class m1_beamlet(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("reserved", C.c_int32), ("beam", m1_beam)]


# (name, restype, argtypes) of every exported function.
SIGNATURES = [
    ("m1_abi_version", C.c_int, []),
    ("m1_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("m1_destroy", None, [C.c_void_p]),
    ("m1_last_error", C.c_char_p, [C.c_void_p]),
    ("m1_set_problem", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                 C.c_double, C.c_double, _ip, C.c_int32,
                                 C.POINTER(m1_material), _dp, C.POINTER(m1_beam)]),
    ("m1_set_state", C.c_int, [C.c_void_p, _dp]),
    ("m1_get_state", C.c_int, [C.c_void_p, _dp]),
    ("m1_get_dose", C.c_int, [C.c_void_p, _dp]),
    ("m1_zero_dose", C.c_int, [C.c_void_p]),
    ("m1_march", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), _dp, C.c_int32,
                           C.POINTER(C.c_int32), _dp, C.c_int32, _dp, _ip]),
    ("m1_spatial_operator", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), C.c_double, _dp, _dp]),
    ("m1_compute_energy_step", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), C.c_double, _dp]),
    ("m1_march_moving", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), C.c_int32, C.c_double,
                                  _dp, C.c_int32, C.POINTER(C.c_int32), _dp, C.c_int32, _dp,
                                  _ip, _ip]),
    ("m1_compute_energy_step_vacuum", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), C.c_double,
                                                C.c_double, _dp]),
    ("m1_gamma_pass_count", C.c_int, [C.c_int32, _dp, _dp, C.c_int32, C.c_int32, C.c_int32, _ip,
                                      C.c_int32, _dp, _dp, C.c_int32, C.c_double,
                                      C.POINTER(C.c_int64)]),
    ("m1_gamma_pass_count_noalloc", C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                              C.c_int32, C.c_int32, C.c_void_p, C.c_int32, _dp,
                                              _dp, C.c_int32, C.c_double, C.c_void_p,
                                              C.POINTER(C.c_int64)]),
    ("m1_gamma_last_error", C.c_char_p, []),
    ("m1_step", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), C.c_int32, C.c_double,
                          C.c_double, _dp]),
    ("m1_set_profiling", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_set_stage_form", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_kernel_count", C.c_int32, []),
    ("m1_kernel_name", C.c_char_p, [C.c_int32]),
    ("m1_tile_shape", None, [C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                             C.POINTER(C.c_int32)]),
    ("m1_kernel_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double)]),
    ("m1_reset_stats", C.c_int, [C.c_void_p]),
    ("m1_tile_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("m1_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("m1_batch_destroy", None, [C.c_void_p]),
    ("m1_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("m1_batch_set_problem", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                       C.c_double, C.c_double, _ip, C.c_int32,
                                       C.POINTER(m1_material), C.c_int32,
                                       C.POINTER(m1_beamlet), _dp]),
    ("m1_batch_march", C.c_int, [C.c_void_p, C.POINTER(m1_numerics), _dp, _ip]),
    ("m1_batch_beamlet_dose", C.c_int, [C.c_void_p, C.c_int32, _dp]),
    ("m1_batch_set_profiling", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_kernel_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double)]),
    ("m1_batch_tile_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]),
    ("m1_batch_reset_stats", C.c_int, [C.c_void_p]),
    ("m1_batch_host_ms", C.c_int, [C.c_void_p, _dp]),
    ("m1_batch_set_cfl_all_tiles", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_set_persistent", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_tile_shape_count", C.c_int32, []),
    ("m1_batch_tile_shape", C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]),
    ("m1_batch_set_tile_shape", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_set_graphs", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_graph_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_char_p)]),
    ("m1_batch_set_skip_model", C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int32,
                                          C.c_int32, C.c_int32, C.c_float]),
    ("m1_batch_set_cfl_tail_floor", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_set_device_steps", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_set_quasi_steady", C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_double]),
    ("m1_batch_set_quasi_steady_tolerance", C.c_int, [C.c_void_p, C.c_double]),
    ("m1_batch_set_progress", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_physics_probe", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         _dp, C.c_double, C.c_int32, _dp]),
    ("m1_batch_set_skip_guard", C.c_int, [C.c_void_p, C.c_int32]),
    ("m1_batch_record", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64]),
    ("m1_batch_samples", C.c_int, [C.c_void_p, _fp, _fp, C.c_int64, C.POINTER(C.c_int64)]),
    ("m1_batch_skip_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      _ip]),
    ("m1_mlp_param_count", C.c_int32, []),
    ("m1_mlp_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("m1_mlp_destroy", None, [C.c_void_p]),
    ("m1_mlp_last_error", C.c_char_p, [C.c_void_p]),
    ("m1_mlp_set_params", C.c_int, [C.c_void_p, _fp]),
    ("m1_mlp_get_params", C.c_int, [C.c_void_p, _fp]),
    ("m1_mlp_forward", C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    ("m1_mlp_loss_grad", C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, _dp, _fp]),
    ("m1_mlp_train", C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                               C.c_uint64, _dp]),
]

_lib = None


def load_library(path: os.PathLike | None = None):
    """Load libm1march.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: the MI355X march library must be built first "
            "(python proton-dose-m1_amd/build_native.py)."
        )
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.m1_abi_version() != ABI_VERSION:
        raise RuntimeError("libm1march.so ABI version mismatch; rebuild the library.")
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    return a.ctypes.data_as(_dp)


_SCHEME = {EnergyScheme.FORWARD_EULER: 0, EnergyScheme.SSPRK3: 1, EnergyScheme.IF_SSPRK3: 2}
_WAVE = {WaveSpeeds.EXACT: 0, WaveSpeeds.UNIT: 1}
_SCAT = {Scattering.OFF: 0, Scattering.ROSSI: 1, Scattering.WENTZEL: 2}
_INTERP = {CoefficientInterpolation.LINEAR: 0, CoefficientInterpolation.PCHIP: 1}
PRECISION_CODE = {Precision.FP64: 0, Precision.FP32_BULK: 1, Precision.FP64_FAST: 2,
                  Precision.FP32_BULK_F64_CLOSURE: 3}
FP64 = 0
FP32_BULK = 1
FP64_FAST = 2
FP32_BULK_F64_CLOSURE = 3


def gamma_pass_count(device, dose, reference, points, offsets, dist_terms, dose_tolerance):
    """Number of points with gamma <= 1 (m1_gamma_pass_count; the arguments
    are those proton_solver.metrics.gamma_pass_rate prepares)."""
    lib = load_library()
    dose = np.ascontiguousarray(dose, dtype=np.float64)
    reference = np.ascontiguousarray(reference, dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.int32)
    offs = np.ascontiguousarray(offsets, dtype=np.float64)
    dt = np.ascontiguousarray(dist_terms, dtype=np.float64)
    nx, ny, nz = dose.shape
    n = C.c_int64(0)
    rc = lib.m1_gamma_pass_count(int(device), _ptr(dose), _ptr(reference), nx, ny, nz, _ptr(pts),
                                 len(pts), _ptr(offs), _ptr(dt), len(dt), float(dose_tolerance),
                                 C.byref(n))
    if rc != M1_OK:
        msg = lib.m1_gamma_last_error().decode()
        raise (ValueError if rc == M1_ERR_VALUE else MarchError)(msg)
    return int(n.value)


GAMMA_NOALLOC_MAX_OFFSETS = 16384  # M1_GAMMA_NOALLOC_MAX_OFFSETS


def gamma_pass_count_noalloc(device, dose_ptr, reference_ptr, shape, points_ptr, n_points,
                             offsets, dist_terms, dose_tolerance, stream=0):
    """m1_gamma_pass_count_noalloc: the pass count of device-resident dose,
    reference ((nx, ny, nz) float64) and points ((n_points, 3) int32) given
    by their device addresses; offsets and dist_terms are host arrays, staged
    in the library's fixed device scratch (nothing is allocated).  stream is a
    hipStream_t handle (0: the default stream)."""
    lib = load_library()
    offs = np.ascontiguousarray(offsets, dtype=np.float64)
    dt = np.ascontiguousarray(dist_terms, dtype=np.float64)
    nx, ny, nz = (int(v) for v in shape)
    n = C.c_int64(0)
    rc = lib.m1_gamma_pass_count_noalloc(int(device), C.c_void_p(int(dose_ptr)),
                                         C.c_void_p(int(reference_ptr)), nx, ny, nz,
                                         C.c_void_p(int(points_ptr)), int(n_points), _ptr(offs),
                                         _ptr(dt), len(dt), float(dose_tolerance),
                                         C.c_void_p(int(stream)), C.byref(n))
    if rc != M1_OK:
        msg = lib.m1_gamma_last_error().decode()
        raise (ValueError if rc == M1_ERR_VALUE else MarchError)(msg)
    return int(n.value)


def numerics_struct(config, precision_code: int | None = None) -> m1_numerics:
    """m1_numerics of a SolverConfig (config.py fields the march reads)."""
    n = config.numerics
    ph = config.physics
    if n.energy_scheme not in _SCHEME:
        raise NotImplementedError(f"Energy scheme {n.energy_scheme} has not been migrated yet.")
    if precision_code is None:
        precision_code = PRECISION_CODE[config.device.precision]
    tile = float(getattr(getattr(config, "device", None), "tile_threshold", 0.0))
    return m1_numerics(_SCHEME[n.energy_scheme], _WAVE[n.wave_speeds], _SCAT[ph.scattering],
                       int(bool(ph.nuclear)), _INTERP[ph.coefficient_interpolation],
                       int(precision_code), float(n.max_energy_step), float(n.cfl),
                       float(n.energy_min), float(n.energy_window_sigma), tile,
                       int(bool(getattr(n, "vacuum_aware_cfl", False))),
                       int(getattr(getattr(config, "device", None), "batch_cfl", "phantom") == "box"))


def _table(E, v, slope, keep):
    E = np.ascontiguousarray(E, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    s = None if slope is None else np.ascontiguousarray(slope, dtype=np.float64)
    keep += [E, v, s]
    return m1_table(len(E), _ptr(E), _ptr(v), _ptr(s))


def material_struct(material, keep: list) -> m1_material:
    from .physics.materials import density, radiation_length
    from .physics.nuclear import load_nuclear_table, nuclear_table_slopes
    from .physics.scattering import load_transport_table, transport_table_slopes
    from .physics.stopping_power import stopping_power_interpolant

    E, v, s = stopping_power_interpolant(material)
    sp = _table(E, v, s, keep)
    tr = _table(*load_transport_table(material), transport_table_slopes(material), keep)
    nu = _table(*load_nuclear_table(material), nuclear_table_slopes(material), keep)
    return m1_material(sp, tr, nu, density(material), radiation_length(material))


class MarchError(RuntimeError):
    """HIP runtime failure or non-finite state reported by the library."""


class Context:
    """One libm1march context (device buffers + stream) bound to a problem."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.m1_create(C.byref(h), int(device))
        if rc != M1_OK:
            raise MarchError(self.lib.m1_last_error(None).decode())
        self.h = h
        self.shape = None

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.m1_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == M1_OK:
            return
        msg = self.lib.m1_last_error(self.h).decode()
        if rc == M1_ERR_VALUE:
            raise ValueError(msg)
        if rc == M1_ERR_NOT_IMPLEMENTED:
            raise NotImplementedError(msg)
        raise MarchError(msg)

    # --- problem -------------------------------------------------------
    def set_problem(self, grid, material_map, inflow, beam):
        nx, ny, nz = (int(s) for s in grid.shape)
        ids = np.ascontiguousarray(np.asarray(material_map.ids), dtype=np.int32)
        if ids.shape != (nx, ny, nz):
            raise ValueError("material ids must have the grid shape")
        spatial = np.ascontiguousarray(np.asarray(inflow.spatial), dtype=np.float64)
        if spatial.shape != (nx, ny):
            raise ValueError("inflow profile must have shape (nx, ny)")
        keep = []
        mats = (m1_material * len(material_map.materials))()
        for i, m in enumerate(material_map.materials):
            mats[i] = material_struct(m, keep)
        b = m1_beam(float(beam.energy), float(beam.energy_spread), float(beam.angular_spread))
        self._check(self.lib.m1_set_problem(self.h, nx, ny, nz, float(grid.dx), float(grid.dy),
                                            float(grid.dz), _ptr(ids), len(mats), mats,
                                            _ptr(spatial), C.byref(b)))
        self.shape = (nx, ny, nz)

    @property
    def n(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    def set_state(self, U=None):
        if U is None:
            self._check(self.lib.m1_set_state(self.h, None))
            return
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (4, *self.shape):
            raise ValueError(f"state has shape {U.shape}, expected {(4, *self.shape)}")
        self._check(self.lib.m1_set_state(self.h, _ptr(U)))

    def get_state(self):
        U = np.empty((4, *self.shape))
        self._check(self.lib.m1_get_state(self.h, _ptr(U)))
        return U

    def get_dose(self):
        d = np.empty(self.shape)
        self._check(self.lib.m1_get_dose(self.h, _ptr(d)))
        return d

    # --- march ---------------------------------------------------------
    def march(self, num: m1_numerics, snapshot_energies=()):
        """(energies, {E: U}) after marching; dose and state stay on device."""
        cap = 1 << 20
        energies = np.empty(cap)
        ne = C.c_int32(0)
        snaps = np.ascontiguousarray([float(s) for s in snapshot_energies], dtype=np.float64)
        ns = len(snaps)
        snap_out = np.empty((ns, 4, *self.shape)) if ns else None
        taken = np.zeros(max(ns, 1), dtype=np.int32)
        self._check(self.lib.m1_march(self.h, C.byref(num), _ptr(energies), cap, C.byref(ne),
                                      _ptr(snaps) if ns else None, ns, _ptr(snap_out),
                                      _ptr(taken)))
        snapshots = {}
        for i in range(ns):
            if taken[i]:
                snapshots[float(snaps[i])] = snap_out[i]
        return energies[: min(ne.value, cap)].copy(), snapshots

    def march_moving(self, num: m1_numerics, guard_cells: int, truncation_sigma: float,
                     snapshot_energies=()):
        """Moving active-domain march: (energies, {E: U}, boxes) with boxes[s] =
        (i0, i1, j0, j1, k0, k1) of step s; dose and the embedded final state
        stay on the device."""
        cap = 1 << 20
        energies = np.empty(cap)
        boxes = np.zeros((cap, 6), dtype=np.int32)
        ne = C.c_int32(0)
        snaps = np.ascontiguousarray([float(s) for s in snapshot_energies], dtype=np.float64)
        ns = len(snaps)
        snap_out = np.empty((ns, 4, *self.shape)) if ns else None
        taken = np.zeros(max(ns, 1), dtype=np.int32)
        self._check(self.lib.m1_march_moving(
            self.h, C.byref(num), int(guard_cells), float(truncation_sigma), _ptr(energies), cap,
            C.byref(ne), _ptr(snaps) if ns else None, ns, _ptr(snap_out), _ptr(taken),
            _ptr(boxes)))
        snapshots = {}
        for i in range(ns):
            if taken[i]:
                snapshots[float(snaps[i])] = snap_out[i]
        k = min(ne.value, cap)
        return energies[:k].copy(), snapshots, boxes[: max(k - 1, 0)].copy()

    def compute_energy_step_vacuum(self, num, energy, relative_tolerance):
        out = C.c_double()
        self._check(self.lib.m1_compute_energy_step_vacuum(
            self.h, C.byref(num), float(energy), float(relative_tolerance), C.byref(out)))
        return out.value

    def spatial_operator(self, num, energy, boundary_state=None):
        b = None
        if boundary_state is not None:
            b = np.ascontiguousarray(boundary_state, dtype=np.float64)
            if b.shape != (4, self.shape[0], self.shape[1]):
                raise ValueError(
                    f"Boundary state has shape {b.shape}, expected {(4, *self.shape[:2])}."
                )
        rhs = np.empty((4, *self.shape))
        self._check(self.lib.m1_spatial_operator(self.h, C.byref(num), float(energy), _ptr(b),
                                                 _ptr(rhs)))
        return rhs

    def compute_energy_step(self, num, energy):
        out = C.c_double()
        self._check(self.lib.m1_compute_energy_step(self.h, C.byref(num), float(energy),
                                                    C.byref(out)))
        return out.value

    def step(self, num, kind, energy, dE, want_edep=True):
        edep = np.empty(self.shape) if want_edep else None
        self._check(self.lib.m1_step(self.h, C.byref(num), int(kind), float(energy), float(dE),
                                     _ptr(edep)))
        return edep

    # --- profiling -------------------------------------------------------
    def set_stage_form(self, split: bool):
        """FP64 stages as split passes (True, the default) or the fused tile
        kernel (False); the results are the same bits."""
        self._check(self.lib.m1_set_stage_form(self.h, int(bool(split))))

    def set_profiling(self, on: bool):
        self._check(self.lib.m1_set_profiling(self.h, int(bool(on))))

    def reset_stats(self):
        self._check(self.lib.m1_reset_stats(self.h))

    def tile_stats(self):
        """(tiles evaluated, tiles a full-domain march evaluates) since reset_stats."""
        a = C.c_int64()
        b = C.c_int64()
        self._check(self.lib.m1_tile_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_stats(self):
        out = {}
        for k in range(self.lib.m1_kernel_count()):
            n = C.c_int64()
            ms = C.c_double()
            self._check(self.lib.m1_kernel_stats(self.h, k, C.byref(n), C.byref(ms)))
            out[self.lib.m1_kernel_name(k).decode()] = (n.value, ms.value)
        return out


def tile_shape():
    """Cells per FP32-bulk workgroup tile (x, y, z) of the loaded library."""
    lib = load_library()
    v = [C.c_int32() for _ in range(3)]
    lib.m1_tile_shape(*(C.byref(x) for x in v))
    return tuple(x.value for x in v)


def batch_tile_shapes():
    """Tile shapes (x, y, z cells) the batched kernels are compiled for; index
    0 is the default."""
    lib = load_library()
    out = []
    for i in range(lib.m1_batch_tile_shape_count()):
        v = [C.c_int32() for _ in range(3)]
        if lib.m1_batch_tile_shape(i, *(C.byref(x) for x in v)) != M1_OK:
            raise MarchError("tile shape query failed")
        out.append(tuple(x.value for x in v))
    return out


class Batch:
    """One libm1march batch: a phantom and its beamlet boxes in device memory."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.m1_batch_create(C.byref(h), int(device))
        if rc != M1_OK:
            raise MarchError(self.lib.m1_batch_last_error(None).decode())
        self.h = h
        self.shape = None
        self.boxes = ()
        self.model = None
        self.tile_shape = batch_tile_shapes()[0]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.m1_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == M1_OK:
            return
        msg = self.lib.m1_batch_last_error(self.h).decode()
        if rc == M1_ERR_VALUE:
            raise ValueError(msg)
        if rc == M1_ERR_NOT_IMPLEMENTED:
            raise NotImplementedError(msg)
        raise MarchError(msg)

    def set_problem(self, ids, spacing, materials, boxes, profiles, beams):
        """ids (nx, ny, nz) int; spacing (dx, dy, dz) m; materials: one per id;
        boxes[b] = (x0, y0, nx_b, ny_b, nz_b); profiles[b] (nx_b, ny_b) inflow
        profile; beams[b] with energy, energy_spread, angular_spread."""
        ids = np.ascontiguousarray(np.asarray(ids), dtype=np.int32)
        if ids.ndim != 3:
            raise ValueError("ids must be a 3-D array")
        if not (len(boxes) == len(profiles) == len(beams)):
            raise ValueError("boxes, profiles and beams must have one entry per beamlet")
        keep = []
        mats = (m1_material * len(materials))()
        for i, m in enumerate(materials):
            mats[i] = material_struct(m, keep)
        bl = (m1_beamlet * len(boxes))()
        flat = []
        for i, (box, prof, beam) in enumerate(zip(boxes, profiles, beams)):
            x0, y0, nx, ny, nz = (int(v) for v in box)
            prof = np.asarray(prof, dtype=np.float64)
            if prof.shape != (nx, ny):
                raise ValueError(f"inflow profile {i} must have shape ({nx}, {ny})")
            bl[i] = m1_beamlet(x0, y0, nx, ny, nz, 0,
                               m1_beam(float(beam.energy), float(beam.energy_spread),
                                       float(beam.angular_spread)))
            flat.append(prof.ravel())
        spatial = np.ascontiguousarray(np.concatenate(flat), dtype=np.float64)
        dx, dy, dz = (float(v) for v in spacing)
        self._check(self.lib.m1_batch_set_problem(
            self.h, *(int(v) for v in ids.shape), dx, dy, dz, _ptr(ids), len(materials), mats,
            len(boxes), bl, _ptr(spatial)))
        self.shape = tuple(int(v) for v in ids.shape)
        self.boxes = tuple(tuple(int(v) for v in b) for b in boxes)

    def march(self, num: m1_numerics, want_dose: bool = True):
        """Returns (phantom dose or None, energy steps per beamlet)."""
        dose = np.empty(self.shape, dtype=np.float64) if want_dose else None
        steps = np.zeros(len(self.boxes), dtype=np.int32)
        self._check(self.lib.m1_batch_march(self.h, C.byref(num), _ptr(dose), _ptr(steps)))
        return dose, steps

    def beamlet_dose(self, b: int):
        if not 0 <= int(b) < len(self.boxes):
            raise ValueError("beamlet index out of range")
        x0, y0, nx, ny, nz = self.boxes[b]
        out = np.empty((nx, ny, nz), dtype=np.float64)
        self._check(self.lib.m1_batch_beamlet_dose(self.h, int(b), _ptr(out)))
        return out

    def set_profiling(self, on: bool):
        self._check(self.lib.m1_batch_set_profiling(self.h, int(bool(on))))

    def reset_stats(self):
        self._check(self.lib.m1_batch_reset_stats(self.h))

    def tile_stats(self):
        a = C.c_int64()
        b = C.c_int64()
        self._check(self.lib.m1_batch_tile_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def host_ms(self):
        v = C.c_double()
        self._check(self.lib.m1_batch_host_ms(self.h, C.byref(v)))
        return v.value

    def kernel_stats(self):
        out = {}
        for k in range(self.lib.m1_kernel_count()):
            n = C.c_int64()
            ms = C.c_double()
            self._check(self.lib.m1_batch_kernel_stats(self.h, k, C.byref(n), C.byref(ms)))
            out[self.lib.m1_kernel_name(k).decode()] = (n.value, ms.value)
        return out

    def set_tile_shape(self, index: int):
        """Use the kernels of tile shape `index` (batch_tile_shapes()); call
        before set_problem."""
        self._check(self.lib.m1_batch_set_tile_shape(self.h, int(index)))
        self.tile_shape = batch_tile_shapes()[int(index)]

    def set_cfl_all_tiles(self, enable: bool):
        """Evaluate the CFL over every tile of the boxes (bit-identical to the
        default live-list form, slower)."""
        self._check(self.lib.m1_batch_set_cfl_all_tiles(self.h, int(bool(enable))))

    def set_persistent(self, enable: bool):
        """Persistent stage kernels with per-XCD tile queues (True) or one
        workgroup per tile (False)."""
        self._check(self.lib.m1_batch_set_persistent(self.h, int(bool(enable))))

    def set_graphs(self, enable: bool):
        """Replay plain steps from captured HIP graphs (True) or launch them
        one by one (False)."""
        self._check(self.lib.m1_batch_set_graphs(self.h, int(bool(enable))))

    def graph_info(self):
        """(steps of the last march replayed from a graph, capture error or '')."""
        n = C.c_int64()
        err = C.c_char_p()
        self._check(self.lib.m1_batch_graph_info(self.h, C.byref(n), C.byref(err)))
        return n.value, (err.value or b"").decode()

    def set_skip_model(self, model, decision=0.5, horizon=8, infer_every=4, judge_every=64,
                       judge_tol=1e-3):
        """Use the tile-skip network `model` (an Mlp, or None to switch it off)."""
        h = None if model is None else model.h
        self._check(self.lib.m1_batch_set_skip_model(self.h, h, float(decision), int(horizon),
                                                     int(infer_every), int(judge_every),
                                                     float(judge_tol)))
        self.model = model  # the library keeps a pointer: keep the model alive

    def set_cfl_tail_floor(self, enable: bool):
        """Sub-threshold cells count with the zero state's speeds in the CFL."""
        self._check(self.lib.m1_batch_set_cfl_tail_floor(self.h, int(bool(enable))))

    def set_device_steps(self, enable: bool):
        """Energies, CFL steps, step scalars and judge decisions on the device:
        steps are enqueued without a host round trip between them."""
        self._check(self.lib.m1_batch_set_device_steps(self.h, int(bool(enable))))

    def set_progress(self, every: int):
        """A status line on stderr every `every` device-stepped energy steps (0: none)."""
        self._check(self.lib.m1_batch_set_progress(self.h, int(every)))

    def set_quasi_steady(self, material_mask: int, iterations: int, pseudo_cfl: float = 0.5):
        """Hold the materials of material_mask through each energy step and
        relax them towards div F = 0 after it (`iterations` sweeps; 0: off)."""
        self._check(self.lib.m1_batch_set_quasi_steady(self.h, int(material_mask),
                                                        int(iterations), float(pseudo_cfl)))

    def set_quasi_steady_tolerance(self, tolerance: float):
        """Stop a beamlet's sweeps after a stage once no quasi-steady cell
        changes by more than tolerance x its peak inflow fluence (0: every
        sweep)."""
        self._check(self.lib.m1_batch_set_quasi_steady_tolerance(self.h, float(tolerance)))

    def physics_probe(self, material: int, kind: int, energies, dE: float = 0.0,
                      scattering: int = 2, interpolation: int = 0):
        """Device values of the step physics of `material` at `energies`:
        kind 0 S(E), 1 transport coefficient, 2 removal coefficient,
        3 scattering depth over [E - dE, E], 4 nuclear depth."""
        e = np.ascontiguousarray(energies, dtype=np.float64)
        out = np.empty_like(e)
        self._check(self.lib.m1_batch_physics_probe(self.h, int(material), int(kind),
                                                    int(scattering), int(interpolation), _ptr(e),
                                                    float(dE), int(e.size), _ptr(out)))
        return out

    def set_skip_guard(self, guard_cells: int):
        """Network candidates: 4 keeps them 4 cells from any cell above the
        threshold (the default), 0 admits every tile the reach rule adds."""
        self._check(self.lib.m1_batch_set_skip_guard(self.h, int(guard_cells)))

    def record(self, enable=True, full_domain=True, horizon=8, capacity=1 << 20):
        """Record (features, label) samples during the following marches."""
        self._check(self.lib.m1_batch_record(self.h, int(bool(enable)), int(bool(full_domain)),
                                             int(horizon), int(capacity)))

    def samples(self):
        """Recorded samples: x (n, 11) float32 and labels y (n,) float32."""
        n = C.c_int64()
        self._check(self.lib.m1_batch_samples(self.h, None, None, 0, C.byref(n)))
        x = np.empty((n.value, Mlp.N_INPUTS), dtype=np.float32)
        y = np.empty(n.value, dtype=np.float32)
        if n.value:
            self._check(self.lib.m1_batch_samples(self.h, _ptr(x), _ptr(y), n.value,
                                                  C.byref(n)))
        return x, y

    def skip_stats(self):
        """(judge checks, rejections, per-beamlet step it was sent back to or -1)."""
        j = C.c_int64()
        r = C.c_int64()
        back = np.full(len(self.boxes), -1, dtype=np.int32)
        self._check(self.lib.m1_batch_skip_stats(self.h, C.byref(j), C.byref(r), _ptr(back)))
        return j.value, r.value, back


class Mlp:
    """The tile-skip network on one device (m1_mlp_*)."""

    N_INPUTS = 11

    def __init__(self, device: int = 0, params=None):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.m1_mlp_create(C.byref(h), int(device))
        if rc != M1_OK:
            raise MarchError(self.lib.m1_mlp_last_error(None).decode())
        self.h = h
        self.n_params = int(self.lib.m1_mlp_param_count())
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.m1_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == M1_OK:
            return
        msg = self.lib.m1_mlp_last_error(self.h).decode()
        if rc == M1_ERR_VALUE:
            raise ValueError(msg)
        raise MarchError(msg)

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.N_INPUTS:
            raise ValueError(f"x must have shape (n, {self.N_INPUTS})")
        return x

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters")
        self._check(self.lib.m1_mlp_set_params(self.h, _ptr(p)))

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.m1_mlp_get_params(self.h, _ptr(p)))
        return p

    def forward(self, x):
        x = self._x(x)
        out = np.empty(len(x), dtype=np.float32)
        self._check(self.lib.m1_mlp_forward(self.h, _ptr(x), len(x), _ptr(out)))
        return out

    def loss_grad(self, x, y):
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_double()
        g = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.m1_mlp_loss_grad(self.h, _ptr(x), _ptr(y), len(x), C.byref(loss),
                                              _ptr(g)))
        return loss.value, g

    def train(self, x, y, epochs=5, batch=4096, lr=1e-3, seed=0):
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.m1_mlp_train(self.h, _ptr(x), _ptr(y), len(x), int(epochs),
                                          int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value
