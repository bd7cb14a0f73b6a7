"""The oracle's moving active-domain march and vacuum-aware CFL, CPU only.

or_energy_march_moving restates _moving_active_energy_march
(energy.py:1262-1494) with the box helpers of moving_active_domain.py:35-253;
or_compute_energy_step_vacuum restates compute_energy_step with
vacuum_relative_tolerance (energy.py:459-660).  They are pinned here by:

  * reduction to the static march: with a guard wider than the grid every
    box is the whole grid, and the march must equal or_energy_march bit for
    bit (local problem, dose scatter and embedding add nothing);
  * truncation: with the default guard and sigma the dose stays within 1e-6
    of the static march (cells below exp(-12.5) of the peak are dropped);
  * the box rule restated independently with NumPy (bounding box of the
    significant cells, widened by the guard and clipped to the grid), checked
    against the box every step of a march was taken on;
  * snapshots and the final state being the box state embedded in the grid;
  * the vacuum-aware CFL reducing to the plain CFL when every cell is active
    or nothing is (scale 0), never giving a smaller step, and matching an
    independent NumPy/SciPy evaluation (binary dilation along each axis) on
    random states.
The reference's published moving-domain results at 100 x 100 x 400 cells
(tests/golden/thesis_golden.json) pin the whole march on the GPU, where it
runs in seconds (tests/test_gpu_moving.py).
"""

import math

import numpy as np
import pytest
from scipy import ndimage

from proton_solver import (Beam, Box, EnergyScheme, Geometry, Material, Numerics, Physics,
                           Scattering, SolverConfig, WaveSpeeds)
from proton_solver.physics.stopping_power import stopping_power


def cfg(shape=(20, 20, 60), energy=60.0, physics=Physics(), objects=(), guard=9, sigma=5.0,
        va=False, scheme=EnergyScheme.IF_SSPRK3, wave=WaveSpeeds.EXACT, beam_sigma=0.002):
    size = tuple(0.001 * s for s in shape)
    return SolverConfig(
        geometry=Geometry(size=size, resolution=shape, objects=list(objects)),
        beam=Beam(n_protons=1e6, energy=energy, energy_spread=1.0,
                  sigma=(beam_sigma, beam_sigma)),
        physics=physics,
        numerics=Numerics(energy_scheme=scheme, wave_speeds=wave, cfl=0.9, max_energy_step=0.5,
                          energy_min=1.0, moving_active_domain=True, active_guard_cells=guard,
                          active_truncation_sigma=sigma, vacuum_aware_cfl=va),
    )


# fits every grid used here (at least 8 x 8 x 13 mm)
HETERO = (Box(-0.004, 0.0, -0.004, 0.004, 0.004, 0.008, Material.BONE_COMPACT),
          Box(0.001, 0.004, -0.004, 0.0, 0.0, 0.012, Material.LUNG))


@pytest.fixture(scope="module")
def water(oracle_lib):
    c = cfg()
    op = oracle_lib.OracleProblem.from_config(c)
    return op, op.march(), op.march_moving()


def test_wide_guard_is_the_static_march(oracle_lib):
    for c in (cfg(shape=(12, 12, 40), energy=40.0),
              cfg(shape=(12, 12, 40), energy=40.0, physics=Physics(scattering=Scattering.WENTZEL,
                                                                   nuclear=True),
                  objects=HETERO, scheme=EnergyScheme.SSPRK3)):
        op = oracle_lib.OracleProblem.from_config(c)
        U, dose, E, _ = op.march()
        Um, dm, Em, _, boxes = op.march_moving(guard=1000)
        assert (boxes == [0, 12, 0, 12, 0, 40]).all()
        assert np.array_equal(E, Em)
        assert np.array_equal(dose, dm)
        assert np.array_equal(U, Um)


def test_truncation_keeps_the_dose(water):
    op, (U, dose, E, _), (Um, dm, Em, _, boxes) = water
    err = np.abs(dm - dose).max() / dose.max()
    assert err < 1e-6
    assert np.array_equal(E, Em)  # the step is capped by max_energy_step throughout
    vol = np.prod(boxes[:, 1::2] - boxes[:, ::2], axis=1)
    assert vol.mean() < 0.7 * np.prod(op.shape)
    assert boxes[:, 4].max() > 0          # the box leaves the entrance plane
    assert (boxes[1:, 5] >= boxes[:-1, 5]).all() or boxes[:, 5].max() == op.shape[2]


def _desired_box(op, U_full, prev, energy, tol, guard):
    """moving_active_domain.py:107-193 restated with NumPy on the embedded state."""
    i0, i1, j0, j1, k0, k1 = prev
    phi = U_full[0, i0:i1, j0:j1, k0:k1]
    scale = phi.max()
    phi_in = None
    if k0 == 0:
        phi_in = op.inflow_state(energy)[0, i0:i1, j0:j1]
        scale = max(scale, phi_in.max())
    if scale == 0.0:
        return None
    act = phi > tol * scale
    if phi_in is not None:
        act[:, :, 0] |= phi_in > tol * scale
    idx = np.nonzero(act)
    if len(idx[0]) == 0:
        return None
    lo = [int(a.min()) for a in idx]
    hi = [int(a.max()) + 1 for a in idx]
    nx, ny, nz = op.shape
    return (max(0, i0 + lo[0] - guard), min(nx, i0 + hi[0] + guard),
            max(0, j0 + lo[1] - guard), min(ny, j0 + hi[1] + guard),
            max(0, k0 + lo[2] - guard), min(nz, k0 + hi[2] + guard))


def test_box_rule_against_numpy(oracle_lib):
    c = cfg(shape=(16, 16, 50), energy=50.0, physics=Physics(scattering=Scattering.WENTZEL,
                                                             nuclear=True),
            objects=HETERO, guard=5, sigma=3.0)
    op = oracle_lib.OracleProblem.from_config(c)
    tol = math.exp(-0.5 * 3.0 ** 2)
    _, _, E, _, boxes = op.march_moving()
    n = len(boxes)
    assert n > 20
    for s in sorted({1, 2, 5, n // 3, n // 2, (2 * n) // 3, n - 1}):
        U_s, _, E_s, _, b_s = op.march_moving(max_steps=s)
        assert np.array_equal(b_s, boxes[:s])
        want = _desired_box(op, U_s, tuple(boxes[s - 1]), E[s], tol, 5)
        assert want is not None
        assert tuple(boxes[s]) == want, (s, boxes[s], want)
        # the state is zero outside the box it lives on
        i0, i1, j0, j1, k0, k1 = boxes[s - 1]
        inside = np.zeros(op.shape, bool)
        inside[i0:i1, j0:j1, k0:k1] = True
        assert not U_s[:, ~inside].any()


def test_snapshots_are_embedded_box_states(oracle_lib):
    c = cfg(shape=(12, 12, 40), energy=40.0, guard=4)
    op = oracle_lib.OracleProblem.from_config(c)
    _, _, E, _, _ = op.march_moving()
    k = len(E) // 2
    targets = (float(E[k]) - 0.123, 500.0)
    U, dose, E2, snaps, boxes = op.march_moving(snapshot_energies=targets)
    assert list(snaps) == [targets[0]]
    s = int(np.nonzero(np.abs(E2 - targets[0]) <= 1e-12)[0][0])
    U_s, *_ = op.march_moving(snapshot_energies=targets, max_steps=s)
    assert np.array_equal(snaps[targets[0]], U_s)


def test_initial_state_is_cut_to_the_initial_box(oracle_lib, water):
    op, _, (Um, dm, Em, _, boxes) = water
    U0 = np.zeros((4, *op.shape))
    i0, i1, j0, j1, k0, k1 = boxes[0]
    U0[0, :, :, k1:] = 1e6       # outside the initial box: dropped
    U0[3, :, :, k1:] = 1e5
    U, d, E, _, b = op.march_moving(U=U0)
    assert np.array_equal(d, dm) and np.array_equal(U, Um) and np.array_equal(b, boxes)


def test_argument_checks(oracle_lib):
    op = oracle_lib.OracleProblem.from_config(cfg(shape=(8, 8, 20), energy=30.0))
    with pytest.raises(ValueError, match="active_guard_cells"):
        op.march_moving(guard=3)
    with pytest.raises(ValueError, match="active_truncation_sigma"):
        op.march_moving(truncation_sigma=0.0)


# ------------------------------------------------------------ vacuum-aware CFL
def _random_state(shape, seed, vacuum=0.6):
    rng = np.random.default_rng(seed)
    phi = 1e10 * rng.random(shape) ** 3
    phi[rng.random(shape) < vacuum] = 0.0
    d = rng.normal(size=(3, *shape))
    d /= np.linalg.norm(d, axis=0)
    J = d * (rng.random(shape) * phi)
    return np.concatenate([phi[None], J])


def _va_step_numpy(O, op, U, energy, tol, c):
    """energy.py:459-660 with NumPy/SciPy: eigenvalues from the oracle, masks by
    binary dilation with a 7-cell line along each axis."""
    lo_hi = [O.eigenvalues(U, a) for a in range(3)]
    a = [np.maximum(np.abs(lo), np.abs(hi)) for lo, hi in lo_hi]
    U_in = op.inflow_state(energy)
    li, hi_ = O.eigenvalues(U_in[:, :, :, None], 2)
    a[2][:, :, 0] = np.maximum(a[2][:, :, 0], np.maximum(np.abs(li[:, :, 0]), np.abs(hi_[:, :, 0])))
    scale = max(float(U[0].max()), float(U_in[0].max()))
    if scale != 0.0:
        act = U[0] > tol * scale
        act[:, :, 0] |= U_in[0] > tol * scale
        for ax in range(3):
            st = np.zeros((7, 7, 7), bool)
            idx = [3, 3, 3]
            for d in range(7):
                idx[ax] = d
                st[tuple(idx)] = True
            m = ndimage.binary_dilation(act, structure=st)
            a[ax] = np.where(m, a[ax], 0.0)
    S = np.vectorize(lambda m: stopping_power(op.materials[m], energy))(op.ids)
    g = op.grid
    inv = a[0] / (S * g.dx) + a[1] / (S * g.dy) + a[2] / (S * g.dz)
    return min(c.numerics.cfl / float(inv.max()), c.numerics.max_energy_step)


@pytest.mark.parametrize("seed", [1, 2])
def test_vacuum_cfl_against_numpy(oracle_lib, seed):
    c = cfg(shape=(9, 11, 13), energy=60.0, objects=HETERO)
    op = oracle_lib.OracleProblem.from_config(c)
    U = _random_state(op.shape, seed)
    for tol in (1e-6, 0.3):
        got = op.compute_energy_step_vacuum(U, 61.0, tol)
        want = _va_step_numpy(oracle_lib, op, U, 61.0, tol, c)
        assert got == pytest.approx(want, rel=1e-12)  # S(E): PCHIP of the package vs oracle
        assert got >= op.compute_energy_step(U, 61.0)


def test_vacuum_cfl_special_cases(oracle_lib):
    c = cfg(shape=(9, 11, 13), energy=60.0, objects=HETERO)
    op = oracle_lib.OracleProblem.from_config(c)
    U = _random_state(op.shape, 3, vacuum=0.0)
    U[0] += 1.0                                   # every cell active at tol 0
    assert op.compute_energy_step_vacuum(U, 61.0, 0.0) == op.compute_energy_step(U, 61.0)
    Z = np.zeros((4, *op.shape))
    # far outside the beam spectrum the inflow underflows to 0: nothing is
    # active and every direction counts (energy.py:515-519)
    e_far = 60.0 + 45.0
    assert op.inflow_state(e_far)[0].max() == 0.0
    assert op.compute_energy_step_vacuum(Z, e_far, 1e-3) == op.compute_energy_step(Z, e_far)
    # a lone active cell in bone: the water cells (smallest stopping power)
    # keep one axis each, so the step is above the all-vacuum bound
    assert op.materials[op.ids[2, 5, 5]] is Material.BONE_COMPACT
    Z[0, 2, 5, 5] = 1.0
    assert op.compute_energy_step_vacuum(Z, e_far, 1e-3) > op.compute_energy_step(Z, e_far)
