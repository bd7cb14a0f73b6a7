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
