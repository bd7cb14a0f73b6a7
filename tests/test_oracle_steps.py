"""Closed-form checks of the oracle's single steps and CFL (CPU only).

For a spatially uniform state the z-min inflow ghosts perturb the first
three cells in each of the three stages, so cells nine or more cells away
from z-min see identical neighbours in every stage: their flux divergence is
exactly zero and one IF-SSPRK3 step reduces to the exact source flow of the
transformed variable W = S(E) U:

    U_new = C(tau(dE)) S(E) U / S(E - dE),
    edep  = dE (q0/6 + q1/6 + 2 q2/3),  q_k = phi_k S_k,

with C = diag(exp(-tau_n), exp(-(tau_n + tau_s)) x3).  This pins the
integrating factor, the W = S U transform and the dose quadrature of
energy.py:878-1130 without using the oracle's own stage formulas.
"""

import math

import numpy as np
import pytest

from proton_solver import (
    Beam, EnergyScheme, Geometry, Material, Numerics, Physics, Scattering, SolverConfig,
)
from proton_solver.physics.nuclear import nuclear_removal_coefficient
from proton_solver.physics.scattering import first_transport_coefficient
from proton_solver.physics.stopping_power import stopping_power


def _problem(O, physics, shape=(5, 6, 16), energy=150.0):
    cfg = SolverConfig(
        geometry=Geometry(size=tuple(0.001 * s for s in shape), resolution=shape),
        beam=Beam(n_protons=1e6, energy=energy, energy_spread=1.0, sigma=(0.002, 0.002)),
        physics=physics,
        numerics=Numerics(energy_scheme=EnergyScheme.IF_SSPRK3, cfl=0.9, max_energy_step=0.5,
                          energy_min=1.0),
    )
    return O.OracleProblem.from_config(cfg)


def _gl_depth(rate, energy, dE):
    mid, hw = energy - 0.5 * dE, 0.5 * dE
    off = hw / math.sqrt(3.0)
    return hw * (rate(mid - off) + rate(mid + off))


def test_if_ssprk3_uniform_state_is_exact_source_flow(oracle_lib):
    O = oracle_lib
    op = _problem(O, Physics(scattering=Scattering.WENTZEL, nuclear=True))
    shape = op.shape
    U = np.zeros((4, *shape))
    U[0], U[1], U[2], U[3] = 1e9, 1e8, -2e8, 6e8
    E, dE = 60.0, 0.4  # inflow spectrum at 60 MeV of a 150 MeV beam is exactly 0
    U_new, edep = op.step(U, E, dE, "if_ssprk3")
    w = Material.WATER
    S = lambda e: stopping_power(w, e)
    tn = _gl_depth(lambda e: nuclear_removal_coefficient(w, e) / S(e), E, dE)
    ts = _gl_depth(lambda e: first_transport_coefficient(w, e) / S(e), E, dE)
    tn2 = _gl_depth(lambda e: nuclear_removal_coefficient(w, e) / S(e), E, dE / 2)
    scale = S(E) / S(E - dE)
    interior = U_new[:, :, :, 9:]
    assert np.allclose(interior[0], math.exp(-tn) * scale * 1e9, rtol=1e-13)
    for c, j in ((1, 1e8), (2, -2e8), (3, 6e8)):
        assert np.allclose(interior[c], math.exp(-(tn + ts)) * scale * j, rtol=1e-13)
    q0 = 1e9 * S(E)
    q1 = math.exp(-tn) * S(E) * 1e9            # phi_1 S_1 = e^{-tau_n(dE)} S(E) phi
    q2 = math.exp(-tn2) * S(E) * 1e9           # stage 2 at dE/2
    expect = dE * (q0 / 6 + q1 / 6 + 2 * q2 / 3)
    assert np.allclose(edep[:, :, 9:], expect, rtol=1e-13)


def test_ssprk3_uniform_state_is_csda_rescaling(oracle_lib):
    O = oracle_lib
    op = _problem(O, Physics())
    U = np.zeros((4, *op.shape))
    U[0], U[3] = 5e9, 4e9
    E, dE = 70.0, 0.3
    U_new, edep = op.step(U, E, dE, "ssprk3")
    S = lambda e: stopping_power(Material.WATER, e)
    assert np.allclose(U_new[:, :, :, 9:], U[:, :, :, 9:] * S(E) / S(E - dE), rtol=1e-14)
    assert np.allclose(edep[:, :, 9:], dE * 5e9 * S(E), rtol=1e-13)


def test_if_ssprk3_without_sources_is_ssprk3(oracle_lib):
    O = oracle_lib
    op = _problem(O, Physics())
    rng = np.random.default_rng(0)
    U = rng.random((4, *op.shape)) * 1e9
    U[1:] *= 0.2
    a = op.step(U, 150.0, 0.25, "if_ssprk3")
    b = op.step(U, 150.0, 0.25, "ssprk3")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_cfl_of_empty_domain_closed_form(oracle_lib):
    """Empty state: isotropic speeds 1/sqrt(3) everywhere; at z = 0 the
    free-streaming inflow raises the z speed to lambda_max(f = 1 - 1e-10)."""
    O = oracle_lib
    op = _problem(O, Physics(), energy=100.0)
    U = np.zeros((4, *op.shape))
    E = 100.0
    S = stopping_power(Material.WATER, E)
    f = 1.0 - 1e-10
    chi = (5 - 2 * math.sqrt(4 - 3 * f * f)) / 3
    den = 5 - 3 * chi
    disc = max(4 * f * f - den * ((3 * chi - 1) - 2 * (1 - chi)), 0.0)
    lam = (2 * f + math.sqrt(disc)) / den
    iso = 1 / math.sqrt(3)
    h = 0.001
    inv = (2 * iso + max(iso, lam)) / (S * h)
    dE = op.compute_energy_step(U, E)
    assert dE == pytest.approx(min(0.9 / inv, 0.5), rel=1e-14)


def test_zero_energy_spread_rejected(oracle_lib):
    O = oracle_lib
    op = _problem(O, Physics())
    op.prob.energy_spread = 0.0
    with pytest.raises(ValueError, match="Zero energy spread"):
        op.inflow_state(100.0)


# ------------------------------------------------ forward Euler and Strang
def _scheme_problem(O, scheme, physics, shape=(5, 6, 16), energy=150.0, energy_min=1.0):
    cfg = SolverConfig(
        geometry=Geometry(size=tuple(0.001 * s for s in shape), resolution=shape),
        beam=Beam(n_protons=1e6, energy=energy, energy_spread=1.0, sigma=(0.002, 0.002)),
        physics=physics,
        numerics=Numerics(energy_scheme=scheme, cfl=0.9, max_energy_step=0.5,
                          energy_min=energy_min),
    )
    return O.OracleProblem.from_config(cfg)


def _uniform(op, phi, jx, jy, jz):
    U = np.zeros((4, *op.shape))
    U[0], U[1], U[2], U[3] = phi, jx, jy, jz
    return U


def test_forward_euler_uniform_state_is_csda_rescaling(oracle_lib):
    """energy.py:663-722 on a uniform state.  One spatial operator call: the
    z-min ghosts reach three cells, so from k = 3 on the divergence is exactly
    zero and U_new = S(E) U / S(E - dE), edep = dE phi S(E)."""
    O = oracle_lib
    op = _scheme_problem(O, EnergyScheme.FORWARD_EULER, Physics())
    U = _uniform(op, 3e9, 2e8, -1e8, 2.5e9)
    E, dE = 70.0, 0.35
    U_new, edep = op.step(U, E, dE, "forward_euler")
    S = lambda e: stopping_power(Material.WATER, e)
    assert np.allclose(U_new[:, :, :, 3:], U[:, :, :, 3:] * S(E) / S(E - dE), rtol=1e-14)
    assert np.allclose(edep[:, :, 3:], dE * 3e9 * S(E), rtol=1e-14)
    # the cells the inflow ghosts reach do change (the empty inflow drains them)
    assert not np.allclose(U_new[:, :, :, 0], U[:, :, :, 0] * S(E) / S(E - dE), rtol=1e-6)
    # the deposition is taken from the old state only: it does not see the flux
    assert np.allclose(edep, dE * 3e9 * S(E), rtol=1e-14)


def test_forward_euler_is_one_explicit_step_of_the_spatial_operator(oracle_lib):
    """W = S U advanced by dE times -div F of the *old* state at the *old*
    energy (the spatial operator is pinned on its own in test_oracle_pins):
    S(E - dE) U_new = S(E) U + dE rhs(U, E) wherever the projection is idle."""
    O = oracle_lib
    op = _scheme_problem(O, EnergyScheme.FORWARD_EULER, Physics())
    rng = np.random.default_rng(5)
    U = np.zeros((4, *op.shape))
    U[0] = 1e9 * (1.0 + rng.random(op.shape))
    U[3] = 0.5 * U[0]
    U[1] = 0.1 * U[0] * rng.uniform(-1, 1, op.shape)
    U[2] = 0.1 * U[0] * rng.uniform(-1, 1, op.shape)
    E, dE = 90.0, 0.05
    U_new, edep = op.step(U, E, dE, "forward_euler")
    rhs = op.spatial_operator(U, E)
    S = lambda e: stopping_power(Material.WATER, e)
    expect = (S(E) * U + dE * rhs) / S(E - dE)
    idle = np.sqrt((expect[1:] ** 2).sum(0)) < (1 - 1e-6) * expect[0]
    assert idle.mean() > 0.9
    for q in range(4):
        assert np.allclose(U_new[q][idle], expect[q][idle], rtol=1e-12, atol=1e-12 * U[0].max())
    assert np.allclose(edep, dE * U[0] * S(E), rtol=1e-14)
    # a step of the wrong sign or at the new energy would differ by O(dE rhs)
    assert np.abs(dE * rhs).max() > 1e-4 * U[0].max()


@pytest.mark.parametrize("scheme, reach", [(EnergyScheme.SSPRK3, 9),
                                           (EnergyScheme.FORWARD_EULER, 3)],
                         ids=["ssprk3", "forward_euler"])
def test_strang_uniform_state_is_half_flows_around_csda(oracle_lib, scheme, reach):
    """energy.py:1194-1260 on a uniform state: the half-step source flows
    exp(-tau/2) (removal on every moment, Wentzel scattering on J only), the
    CSDA step (a pure S(E)/S(E - dE) rescale away from the inflow ghosts),
    then the same half flows.  tau from the Gauss-Legendre closed form; the
    deposition is the transport step's, taken on the half-stepped state."""
    O = oracle_lib
    op = _scheme_problem(O, scheme, Physics(scattering=Scattering.WENTZEL, nuclear=True))
    U = _uniform(op, 2e9, -3e8, 1e8, 1.2e9)
    E, dE = 60.0, 0.4
    U_new, edep = op.step(U, E, dE, "strang_step")
    w = Material.WATER
    S = lambda e: stopping_power(w, e)
    tn = _gl_depth(lambda e: nuclear_removal_coefficient(w, e) / S(e), E, dE)
    ts = _gl_depth(lambda e: first_transport_coefficient(w, e) / S(e), E, dE)
    scale = S(E) / S(E - dE)
    inner = U_new[:, :, :, reach:]
    assert np.allclose(inner[0], math.exp(-tn) * scale * 2e9, rtol=1e-13)
    for c, j in ((1, -3e8), (2, 1e8), (3, 1.2e9)):
        assert np.allclose(inner[c], math.exp(-(tn + ts)) * scale * j, rtol=1e-13)
    assert np.allclose(edep[:, :, reach:], dE * math.exp(-0.5 * tn) * 2e9 * S(E), rtol=1e-13)


def test_strang_without_sources_is_the_transport_step(oracle_lib):
    O = oracle_lib
    op = _scheme_problem(O, EnergyScheme.SSPRK3, Physics())
    rng = np.random.default_rng(2)
    U = rng.random((4, *op.shape)) * 1e9
    U[1:] *= 0.2
    a = op.step(U, 120.0, 0.25, "strang_step")
    b = op.step(U, 120.0, 0.25, "ssprk3")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# -------------------------------------------------------- snapshot landing
def _march_problem(O, energy_min, scheme=EnergyScheme.IF_SSPRK3):
    return _scheme_problem(O, scheme, Physics(scattering=Scattering.WENTZEL, nuclear=True),
                           shape=(4, 4, 24), energy=40.0, energy_min=energy_min)


@pytest.mark.parametrize("scheme", [EnergyScheme.IF_SSPRK3, EnergyScheme.SSPRK3],
                         ids=["if_ssprk3", "strang"])
def test_snapshot_is_the_march_stopped_at_that_energy(oracle_lib, scheme):
    """energy.py:1567-1578, 1627-1637: the step that would cross a requested
    energy E* is shortened to land on it, and the state there is recorded.
    The march stopped at energy_min = E* takes the same steps (its last step
    is shortened the same way, energy.py:1580-1584), so the snapshot must be
    its final state bit for bit, E* must be one of the energies, and the
    energies up to E* must be the same."""
    O = oracle_lib
    target = 27.3
    U_s, dose_s, E_s, snaps = _march_problem(O, 1.0, scheme).march(snapshot_energies=[target])
    U_t, dose_t, E_t, _ = _march_problem(O, target, scheme).march()
    assert list(snaps) == [target]
    assert target in E_s.tolist() and E_t[-1] == target
    k = E_s.tolist().index(target)
    assert np.array_equal(E_s[: k + 1], E_t)
    assert np.array_equal(snaps[target], U_t)
    assert E_s[-1] == 1.0 and k < len(E_s) - 1
    # landing moved one step boundary: the grids differ right after E*
    plain = _march_problem(O, 1.0, scheme).march()[2]
    assert target not in plain.tolist()
    assert np.array_equal(plain[:k], E_s[:k])


def test_snapshot_targets_sorted_deduplicated_and_bounded(oracle_lib):
    """Requested energies outside [energy_min, E0] are ignored, duplicates
    recorded once, and targets are met in descending order whatever the
    request order: with {21.5, 33.0 (twice), 80 (above E0), 0.5 (below
    energy_min)} the march lands on 33.0 then 21.5, and the state at 21.5 is
    that of the march stopped at 21.5 that lands on 33.0 on its way."""
    O = oracle_lib
    req = [21.5, 33.0, 80.0, 33.0, 0.5]
    _, _, E_s, snaps = _march_problem(O, 1.0).march(snapshot_energies=req)
    assert sorted(snaps) == [21.5, 33.0]
    E_list = E_s.tolist()
    assert E_list.index(33.0) < E_list.index(21.5)
    U_t, _, E_t, snaps_t = _march_problem(O, 21.5).march(snapshot_energies=[33.0])
    assert np.array_equal(snaps[21.5], U_t)
    assert np.array_equal(snaps[33.0], snaps_t[33.0])
    assert np.array_equal(E_s[: len(E_t)], E_t)
