"""Pins the oracle's strain energy, total potential and energy line search
(north-star additions with no counterpart in the reference, which evaluates
no energy and runs only the projection rule of
src/simulation/nlsolver/LineSearchFEA.h:37-68).

  * dW/dF = P (the reference's stress, or_piola) by central differences,
    W(I) = 0, objectivity; the device's W (efem_material_point) agrees;
  * grad of the total potential = the residual (central differences) on
    conservative models (gravity, nodal forces), NaN with pressure faces;
  * the energy rule on a one-dimensional quadratic a x^2 / 2 - b x from 0
    with d = b, whose closed form separates it from the projection rule:
    step 1 passes |g(1) d| <= |g(0) d| iff a <= 2 and the Armijo decrease
    with c1 = 1e-4 iff a <= 2 (1 - c1); a = 1.99995 passes the first and
    fails the second (it would pass with c1 = 0), so the energy rule takes a
    second trial, step 0.5; a = 1.9997 passes both;
  * both rules reach A^-1 b on SPD quadratics, and the same equilibrium on
    a self-weight cantilever.
"""
import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import smooth_u

LAWS = [(0, [114000.0, 0.0, 0.0, 1e-6]), (1, [80000.0, 20000.0, 0.0, 1e-6]), (2, [5e5, -1e4, 1e3, 1e-7])]


@pytest.mark.parametrize("law,p", LAWS)
def test_energy_density_derivative_is_the_reference_stress(law, p, rng):
    assert abs(O.energy_density(law, p, np.eye(3))) < 1e-9 * p[0]
    for _ in range(6):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.5:
            continue
        P = O.piola(law, p, F)
        h = 1e-6
        dW = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[i, j] += h
                Fm[i, j] -= h
                dW[i, j] = (O.energy_density(law, p, Fp) - O.energy_density(law, p, Fm)) / (2 * h)
        assert np.abs(dW - P).max() <= 1e-6 * np.abs(P).max() + 1e-3
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        Q *= np.sign(np.linalg.det(Q))
        W = O.energy_density(law, p, F)
        assert abs(O.energy_density(law, p, Q @ F) - W) <= 1e-9 * abs(W) + 1e-9
        # the device code's energy (cancellation-free form) is the same function
        _, Wd, _ = E.material_point(law, p, F)
        assert abs(Wd - W) <= 1e-9 * abs(W) + 1e-6


@pytest.mark.parametrize("kind", [8, 10, 4])
def test_potential_gradient_is_the_residual(kind, rng):
    m0 = E.Mesh.workload(0, [4, 2, 2, 0.004, 0.002, 0.002, kind, 1, 500.0, 0.05 if kind in (4, 10) else 0.0,
                             3, 0])
    top = np.unique(m0.faces.ravel())
    m = E.Mesh(m0.X, m0.conn, m0.law, m0.params, m0.fixed, None, 0.0, 0, fnodes=top[:4],
               fvec=np.array([0.3, -0.1, 0.2]), gravity=6, consistent_inertia=1, density=[1000.0])
    om = O.Model(m)
    u = smooth_u(m, rng, 0.02)
    v = smooth_u(m, rng, 1.0)
    # fourth-order central difference (the near-incompressible volumetric
    # term makes the potential stiff along v)
    h = 1e-6
    P = om.potential
    fd = (8 * (P(u + h * v) - P(u - h * v)) - (P(u + 2 * h * v) - P(u - 2 * h * v))) / (12 * h)
    g = om.residual(u) @ v
    assert abs(fd - g) <= 1e-6 * abs(g)
    assert abs(om.potential(np.zeros(m.ndof))) < 1e-18
    assert np.isnan(O.Model(m0).potential(u))


def test_energy_rule_rejects_a_step_the_projection_rule_accepts():
    a, b = 1.99995, 1.0
    resid = lambda x: a * x - b
    energy = lambda x: float(0.5 * a * x[0] ** 2 - b * x[0])
    _, stp = O.lbfgs(1, resid, np.zeros(1), epsilon=0.0, max_iterations=1, trace_cap=1)
    _, ste = O.lbfgs(1, resid, np.zeros(1), epsilon=0.0, max_iterations=1, trace_cap=1, energy=energy)
    assert stp["trace"][0, 1:].tolist() == [1.0, 1.0]   # projection: step 1, one trial
    assert ste["trace"][0, 1:].tolist() == [0.5, 2.0]   # energy: step 0.5, two trials
    # with a = 1.9997 <= 2 (1 - c1) step 1 passes both conditions
    a = 1.9997
    _, ste = O.lbfgs(1, lambda x: a * x - b, np.zeros(1), epsilon=0.0, max_iterations=1, trace_cap=1,
                     energy=lambda x: float(0.5 * a * x[0] ** 2 - b * x[0]))
    assert ste["trace"][0, 1:].tolist() == [1.0, 1.0]


@pytest.mark.parametrize("use_h0", [False, True])
def test_energy_rule_solves_spd_quadratics(use_h0, rng):
    n = 30
    A = rng.standard_normal((n, n))
    A = A @ A.T + 0.1 * np.eye(n)
    b = rng.standard_normal(n)
    M = np.diag(np.diag(A))
    h0 = (lambda v: np.linalg.solve(M, v)) if use_h0 else None
    x, st = O.lbfgs(n, lambda x: A @ x - b, np.zeros(n), h0=h0, epsilon=1e-12, max_iterations=5000,
                    energy=lambda x: float(0.5 * x @ A @ x - b @ x))
    # the stopping test is |r|_inf <= eps or <= eps |x|_inf (LBFGS.h:153)
    assert st["rinf"] <= 1e-12 * max(1.0, np.abs(x).max())
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-8 * np.abs(np.linalg.solve(A, b)).max())


def test_energy_rule_reaches_the_projection_equilibrium_under_self_weight():
    m0 = E.Mesh.workload(1, [10, 2, 2, 0, 100.0])
    m = E.Mesh(m0.X, m0.conn, m0.law, m0.params, m0.fixed, None, 0.0, m0.reduced, gravity=6,
               consistent_inertia=1, density=[1050.0])
    om = O.Model(m)
    up, stp = om.solve(epsilon=1e-9, max_iterations=5000, factor_scalar=True)
    ue, ste = om.solve(epsilon=1e-9, max_iterations=5000, factor_scalar=True, line_search=1)
    assert stp["rinf"] <= 1e-9 and ste["rinf"] <= 1e-9
    assert np.linalg.norm(ue - up) <= 1e-6 * np.linalg.norm(up)
    # and the potential at the equilibrium is below the start's
    assert om.potential(ue) < om.potential(np.zeros(m.ndof))
    with pytest.raises(ValueError):
        O.Model(m0).solve(max_iterations=5, line_search=1)


def test_potential_difference_equals_difference_of_totals(rng):
    # Pi(x) - Pi(xp) element by element equals the difference of totals,
    # and its rounding scale is the size of the energy terms at xp
    m0 = E.Mesh.workload(1, [10, 2, 2, 0, 100.0])
    m = E.Mesh(m0.X, m0.conn, m0.law, m0.params, m0.fixed, None, 0.0, 0, gravity=6, consistent_inertia=1,
               density=[1050.0])
    om = O.Model(m)
    xp = smooth_u(m, rng, 0.02)
    x = smooth_u(m, rng, 0.03)
    d, scale = om.potential_diff(x, xp)
    assert abs(d - (om.potential(x) - om.potential(xp))) <= 1e-10 * scale
    assert scale >= abs(om.potential(xp))
