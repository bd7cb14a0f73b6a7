"""Pins of the CPU oracle (oracle/liboracle_m1.so) against mathematics and libraries.

Each reference function the oracle restates is checked against something
that fixes its result independently of the oracle's own formula: SciPy's
PCHIP, numpy.interp, exactness of quadratures and reconstructions on
polynomials, eigenvalues of a numerically differentiated flux Jacobian,
textbook closure formulas, and closed-form single steps.  The whole march is
pinned against the reference's own published FP64 results in
test_golden_march.py.  CPU only.
"""

import math

import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator
from scipy.special import erf

from proton_solver import Material
from proton_solver.physics.nuclear import load_nuclear_table
from proton_solver.physics.scattering import load_transport_table
from proton_solver.physics.stopping_power import load_stopping_power_table

MATERIALS = list(Material)


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# ---------------------------------------------------------------- tables
@pytest.mark.parametrize("material", MATERIALS, ids=lambda m: m.name)
def test_pchip_stopping_power_matches_scipy(O, material):
    E, S = load_stopping_power_table(material)
    ref = PchipInterpolator(E, S)
    rng = np.random.default_rng(0)
    x = np.exp(rng.uniform(np.log(E[0]), np.log(E[-1]), 4000))
    x = np.concatenate([x, E[:50], E[-50:], [E[0], E[-1]]])
    got = np.array([O.table_eval(E, S, v, "PCHIP") for v in x])
    assert np.max(np.abs(got / ref(x) - 1)) < 1e-13
    slopes = O.pchip_slopes(E, S)
    dref = ref.derivative()(E)
    assert np.max(np.abs(slopes - dref) / np.abs(dref).max()) < 1e-12


@pytest.mark.parametrize("material", MATERIALS, ids=lambda m: m.name)
def test_linear_coefficients_match_numpy_interp(O, material):
    rng = np.random.default_rng(1)
    for E, v in (load_transport_table(material), load_nuclear_table(material)):
        x = np.concatenate([rng.uniform(E[0], E[-1], 2000), E[::97], [E[0], E[-1]]])
        got = np.array([O.table_eval(E, v, e, "LINEAR") for e in x])
        assert np.array_equal(got, np.interp(x, E, v))


def test_table_range_is_checked_with_isclose(O):
    E, S = load_stopping_power_table(Material.WATER)
    # 1e-4 MeV lies a few ulps below the first node and is accepted (clamped).
    assert O.table_eval(E, S, 1e-4) == S[0]
    assert O.table_eval(E, S, E[-1]) == S[-1]
    with pytest.raises(ValueError, match="outside table range"):
        O.table_eval(E, S, 0.5e-4)
    with pytest.raises(ValueError, match="outside table range"):
        O.table_eval(E, S, 301.0)


def test_pchip_reproduces_linear_data_and_preserves_monotonicity(O):
    E = np.linspace(1.0, 10.0, 11)
    v = 3.0 + 2.0 * E
    for x in np.linspace(1.0, 10.0, 57):
        assert O.table_eval(E, v, x) == pytest.approx(3.0 + 2.0 * x, rel=1e-14)
    # monotone data with a flat section: interpolant never overshoots
    v2 = np.array([0, 0, 0, 1, 5, 5, 5, 6, 9, 9, 9], dtype=float)
    xs = np.linspace(1.0, 10.0, 901)
    vals = np.array([O.table_eval(E, v2, x) for x in xs])
    assert np.all(np.diff(vals) >= -4 * np.finfo(float).eps * 9.0)


def _synthetic_material(O, S_const=500.0, sig=(2.0, 0.03), nu=(1.0, 0.01), X0=0.36):
    """Material with constant S and linear Sigma(E) tables."""
    E = np.array([0.0001, 300.0])
    return O.material_from_arrays((E, np.full(2, S_const)),
                                  (E, sig[0] + sig[1] * E),
                                  (E, nu[0] + nu[1] * E), 1000.0, X0)


def test_gauss_legendre_depths_exact_for_linear_rates(O):
    mat, keep = _synthetic_material(O)
    for energy, dE in ((150.0, 0.25), (80.0, 3.0), (5.0, 4.9)):
        lo = energy - dE
        exact_n = ((1.0 * (energy - lo)) + 0.01 * (energy**2 - lo**2) / 2) / 500.0
        exact_s = ((2.0 * (energy - lo)) + 0.03 * (energy**2 - lo**2) / 2) / 500.0
        assert O.nuclear_depth(mat, energy, dE) == pytest.approx(exact_n, rel=1e-13)
        assert O.scattering_depth(mat, "WENTZEL", energy, dE) == pytest.approx(exact_s, rel=1e-13)
    with pytest.raises(ValueError, match="dE must be positive"):
        O.nuclear_depth(mat, 10.0, 0.0)


def test_gauss_legendre_depth_matches_quadrature_on_real_tables(O):
    from scipy.integrate import quad

    mat, keep = O.material_struct(Material.WATER)
    E, S = load_stopping_power_table(Material.WATER)
    Sp = PchipInterpolator(E, S)
    Et, St = load_nuclear_table(Material.WATER)
    for energy, dE in ((150.0, 0.25), (60.0, 0.1)):
        ref, _ = quad(lambda e: np.interp(e, Et, St) / Sp(e), energy - dE, energy,
                      epsabs=0, epsrel=1e-13, limit=200)
        # the linear-interpolated table has kinks inside the interval, so the
        # two-point rule is only close here (it is exact for linear rates above)
        assert O.nuclear_depth(mat, energy, dE) == pytest.approx(ref, rel=1e-5)


def test_rossi_coefficient_closed_form(O):
    mat, keep = _synthetic_material(O, X0=0.36)
    energy, dE = 100.0, 1e-3
    m = 938.27208816
    pv = lambda e: e * (e + 2 * m) / (e + m)
    rate = lambda e: (15.0 / pv(e)) ** 2 / 0.36 / 500.0
    from scipy.integrate import quad

    ref, _ = quad(rate, energy - dE, energy, epsrel=1e-14)
    assert O.scattering_depth(mat, "ROSSI", energy, dE) == pytest.approx(ref, rel=1e-9)


# ------------------------------------------------------------- WENO-Z
def test_weno_exact_for_constant_and_linear_data(O):
    assert O.weno_z_left(7.0, 7.0, 7.0, 7.0, 7.0) == 7.0
    for a, b in ((0.0, 1.0), (3.5, -2.25), (1e10, 3e8)):
        v = [a + b * i for i in range(5)]
        # face between cells 2 and 3 sits at 2.5 in cell units
        assert O.weno_z_left(*v) == pytest.approx(a + 2.5 * b, rel=1e-14, abs=1e-14 * abs(b))


def test_weno_linear_weights_give_fifth_order_combination(O):
    # With equal smoothness indicators the weights are the ideal (0.1, 0.6, 0.3)
    # and the face value is (2 v0 - 13 v1 + 47 v2 + 27 v3 - 3 v4) / 60.
    v = [1.0, 2.0, 3.0, 4.0, 5.0]
    assert O.weno_z_left(*v) == pytest.approx((2 - 26 + 141 + 108 - 15) / 60.0, rel=1e-15)


def test_weno_fifth_order_convergence_on_smooth_data(O):
    errs = []
    for n in (16, 32, 64, 128):
        h = 1.0 / n
        x0 = 0.3
        faces = []
        for k in range(5):
            a = x0 + (k - 2) * h - h / 2
            faces.append((math.cos(a) - math.cos(a + h)) / h)  # cell average of sin
        errs.append(abs(O.weno_z_left(*faces) - math.sin(x0 + h / 2)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(orders) > 4.5


def test_weno_is_essentially_non_oscillatory_at_a_jump(O):
    left = O.weno_z_left(0.0, 0.0, 0.0, 1.0, 1.0)
    assert abs(left) < 1e-6
    right = O.weno_z_left(1.0, 1.0, 1.0, 0.0, 0.0)
    assert abs(right - 1.0) < 1e-6


def test_weno_is_odd_and_translation_invariant(O):
    rng = np.random.default_rng(5)
    for _ in range(100):
        v = rng.normal(size=5) * 10
        assert O.weno_z_left(*(-v)) == -O.weno_z_left(*v)


def test_reconstruction_uses_mirrored_stencil_for_right_states(O):
    rng = np.random.default_rng(6)
    U = rng.random((4, 3, 4, 5))
    Ue = O.fill_ghost_cells(U, rng.random((4, 3, 4)))
    Ue[1:] = 0.0  # no face projection acts on J = 0, phi >= 0
    # interior cell (1, 1, 2) sits at (4, 4, 5) of the ghosted array
    lines = {0: Ue[0, :, 4, 5], 1: Ue[0, 4, :, 5], 2: Ue[0, 4, 4, :]}
    for axis, line in lines.items():
        L, R = O.reconstruct(Ue, axis)
        assert L.shape[axis + 1] == U.shape[axis + 1] + 1
        for f in range(U.shape[axis + 1] + 1):
            idx = [1, 1, 2]
            idx[axis] = f
            assert L[(0, *idx)] == O.weno_z_left(*line[f:f + 5])
            assert R[(0, *idx)] == O.weno_z_left(*line[f + 1:f + 6][::-1])


# ------------------------------------------------------- ghost cells
def test_zmin_ghosts_exact_for_quartics(O):
    h = 1.0
    rng = np.random.default_rng(2)
    c = rng.normal(size=5)
    P = np.polynomial.Polynomial(c)
    I = P.integ()

    def avg(a, b):
        return (I(b) - I(a)) / (b - a)

    nz = 6
    U = np.zeros((4, 1, 1, nz))
    for k in range(nz):
        U[:, 0, 0, k] = avg(k * h, (k + 1) * h)
    U_in = np.full((4, 1, 1), P(0.0))
    Ue = O.fill_ghost_cells(U, U_in)
    for g in (1, 2, 3):
        exact = avg(-g * h, -(g - 1) * h)
        assert Ue[0, 3, 3, 3 - g] == pytest.approx(exact, rel=1e-11, abs=1e-11)


def test_ghosts_zero_gradient_elsewhere(O):
    rng = np.random.default_rng(3)
    U = rng.random((4, 3, 4, 6))
    Ue = O.fill_ghost_cells(U, rng.random((4, 3, 4)))
    core = Ue[:, 3:-3, 3:-3, 3:-3]
    assert np.array_equal(core, U)
    assert np.array_equal(Ue[:, 3:-3, 3:-3, -3:], np.repeat(U[..., -1:], 3, axis=3))
    assert np.array_equal(Ue[:, :3, 3:-3, 3:-3], np.repeat(U[:, :1], 3, axis=1))
    assert np.array_equal(Ue[:, -3:, 3:-3, 3:-3], np.repeat(U[:, -1:], 3, axis=1))
    assert np.array_equal(Ue[:, 3:-3, :3, 3:-3], np.repeat(U[:, :, :1], 3, axis=2))
    with pytest.raises(ValueError, match="four z cells"):
        O.fill_ghost_cells(U[..., :3], rng.random((4, 3, 4)))


# ------------------------------------------------- closure and HLL
def _levermore_flux(U, axis):
    """Textbook M1 flux: (J_axis, P[axis, :]) with P = phi((1-chi)/2 I + (3chi-1)/2 n n)."""
    phi, J = U[0], U[1:]
    jn = np.linalg.norm(J)
    f = min(max(jn / phi, 0.0), 1.0) if phi > 0 else 0.0
    chi = (5 - 2 * math.sqrt(4 - 3 * f * f)) / 3
    n = J / jn if jn > 0 else np.zeros(3)
    P = phi * ((1 - chi) / 2 * np.eye(3) + (3 * chi - 1) / 2 * np.outer(n, n))
    return np.concatenate([[J[axis]], P[axis]])


def _states(seed, n=40):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        phi = 10 ** rng.uniform(-3, 3)
        f = rng.uniform(0.02, 0.97)
        out.append(np.concatenate([[phi], f * phi * d]))
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_hll_consistency_equals_textbook_flux(O, axis):
    for U in _states(10 + axis):
        F = O.hll_flux(U.reshape(4, 1), U.reshape(4, 1), axis)[:, 0]
        assert np.allclose(F, _levermore_flux(U, axis), rtol=1e-13, atol=1e-13 * U[0])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_eigenvalues_bound_the_flux_jacobian_spectrum(O, axis):
    for U in _states(20 + axis, 25):
        def F(u):
            return O.hll_flux(u.reshape(4, 1), u.reshape(4, 1), axis)[:, 0]

        Jac = np.zeros((4, 4))
        for k in range(4):
            h = 1e-6 * U[0]
            e = np.zeros(4)
            e[k] = h
            Jac[:, k] = (F(U + e) - F(U - e)) / (2 * h)
        ev = np.linalg.eigvals(Jac).real
        lo, hi = O.eigenvalues(U.reshape(4, 1), axis)
        assert lo[0] == pytest.approx(ev.min(), abs=2e-6)
        assert hi[0] == pytest.approx(ev.max(), abs=2e-6)


def test_eigenvalue_limits(O):
    iso = 1 / math.sqrt(3)
    lo, hi = O.eigenvalues(np.array([[1.0], [0.0], [0.0], [0.0]]), 0)
    assert (lo[0], hi[0]) == (-iso, iso)
    lo, hi = O.eigenvalues(np.array([[0.0], [0.0], [0.0], [0.0]]), 2)
    assert (lo[0], hi[0]) == (-iso, iso)
    # free streaming along z: lambda_max -> 1 along z, -> 0 across
    U = np.array([[1.0], [0.0], [0.0], [1.0]])
    lo, hi = O.eigenvalues(U, 2)
    assert hi[0] == pytest.approx(1.0, abs=1e-4)
    lo, hi = O.eigenvalues(U, 0)
    assert abs(hi[0]) < 1e-4 and abs(lo[0]) < 1e-4


def test_unit_wave_speeds_give_rusanov_flux(O):
    rng = np.random.default_rng(4)
    for UL, UR in zip(_states(30, 10), _states(31, 10)):
        F = O.hll_flux(UL.reshape(4, 1), UR.reshape(4, 1), 1, 1)[:, 0]
        ref = 0.5 * (_levermore_flux(UL, 1) + _levermore_flux(UR, 1)) - 0.5 * (UR - UL)
        assert np.allclose(F, ref, rtol=1e-13, atol=1e-13)


def test_closure_limits_through_flux(O):
    F = O.hll_flux(np.array([[2.0], [0], [0], [0]]), np.array([[2.0], [0], [0], [0]]), 0)[:, 0]
    assert F[1] == pytest.approx(2.0 / 3.0, rel=1e-15)  # chi(0) = 1/3
    U = np.array([[2.0], [0.0], [0.0], [2.0]])
    F = O.hll_flux(U, U, 2)[:, 0]
    assert F[3] == pytest.approx(2.0, rel=1e-15)  # chi(1) = 1: P_zz = phi


# ------------------------------------------------------- projection
def test_projection_is_idempotent_and_realizable(O):
    rng = np.random.default_rng(7)
    U = rng.normal(size=(4, 500)) * 10
    P1 = O.project(U)
    P2 = O.project(P1)
    # a rescaled |J| can land one ulp above (1 - eps) phi, so a second pass
    # may rescale by 1 - O(ulp) again
    assert np.allclose(P1, P2, rtol=4e-16, atol=0)
    phi = P1[0]
    jn = np.linalg.norm(P1[1:], axis=0)
    assert np.all(phi >= 0)
    assert np.all(jn <= (1 - 1e-10) * phi * (1 + 1e-15) + 1e-300)
    # direction kept where J was rescaled
    moved = np.linalg.norm(U[1:], axis=0) > 0
    cos = np.sum(U[1:] * P1[1:], axis=0)
    assert np.all(cos[moved & (jn > 0)] > 0)


# ---------------------------------------------------- RK boundary traces
def test_rk3_trace_combination_matches_taylor_for_cubics(O):
    rng = np.random.default_rng(8)
    for dE in (0.25, 1.0, 3.0):
        a, b, c, d = rng.normal(size=4)
        B = lambda t: a + b * t + c * t * t + d * t**3
        s1, s2 = O.rk3_trace_combination(*[np.array([B(f * dE)]) for f in (0, 1 / 3, 2 / 3, 1)])
        assert s1[0] == pytest.approx(a + dE * b, rel=1e-12, abs=1e-12)
        assert s2[0] == pytest.approx(a + 0.5 * dE * b + 0.25 * dE**2 * (2 * c), rel=1e-12, abs=1e-12)


# ----------------------------------------------------- inflow profile
def test_inflow_profile_is_erf_cell_average_with_exact_mass():
    from proton_solver import Beam, Geometry
    from proton_solver.numerics.boundary import BeamInflow
    from proton_solver.numerics.grid import Grid

    geo = Geometry(size=(0.02, 0.03, 0.01), resolution=(20, 30, 10))
    beam = Beam(n_protons=1e6, energy=100, sigma=(0.002, 0.003), center=(0.001, -0.002))
    grid = Grid.from_geometry(geo, np)
    inflow = BeamInflow.from_beam(beam, grid, np)
    assert inflow.spatial.sum() * grid.dx * grid.dy == pytest.approx(1e6, rel=1e-14)
    gx = 0.5 * (erf((grid.x + grid.dx / 2 - 0.001) / (math.sqrt(2) * 0.002))
                - erf((grid.x - grid.dx / 2 - 0.001) / (math.sqrt(2) * 0.002)))
    gy = 0.5 * (erf((grid.y + grid.dy / 2 + 0.002) / (math.sqrt(2) * 0.003))
                - erf((grid.y - grid.dy / 2 + 0.002) / (math.sqrt(2) * 0.003)))
    r2 = ((grid.x[:, None] - 0.001) / 0.002) ** 2 + ((grid.y[None, :] + 0.002) / 0.003) ** 2
    prof = np.where(r2 <= 25.0, np.outer(gx, gy), 0.0)
    ratio = inflow.spatial[prof > 0] / prof[prof > 0]
    assert np.allclose(ratio, ratio[0], rtol=1e-12)
    assert np.all(inflow.spatial[prof == 0] == 0)
