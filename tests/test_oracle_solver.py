"""Pins the oracle's L-BFGS pieces: the two-loop recursion against the
explicit BFGS inverse update (Nocedal & Wright Alg. 7.4 / eq. 7.19), the
projection line search against a hand-enumerated trajectory, and the stopping
rules.

Reference: src/simulation/nlsolver/{LBFGS.h,LBFGSMat.h,LineSearchFEA.h}.
"""
import numpy as np
import pytest

import efem_amd as E
import oracle as O


def dense_bfgs_inverse(S, Y, H0):
    # H_{k+1} = (I - rho s y^T) H_k (I - rho y s^T) + rho s s^T, oldest first
    H = H0.copy()
    n = H.shape[0]
    for s, y in zip(S, Y):
        rho = 1.0 / (s @ y)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H


@pytest.mark.parametrize("use_h0", [False, True])
def test_two_loop_equals_dense_bfgs(use_h0, rng):
    n, mcap, ncorr = 12, 6, 4
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    S = np.zeros((mcap, n))
    Y = np.zeros((mcap, n))
    ys = np.zeros(mcap)
    ptr = mcap
    hist = []
    for _ in range(ncorr + 3):  # wrap the ring buffer
        s = rng.standard_normal(n)
        y = A @ s
        loc = ptr % mcap
        S[loc], Y[loc], ys[loc] = s, y, s @ y
        ptr = loc + 1
        hist.append((s, y))
    nc = min(len(hist), mcap)
    hist = hist[-nc:]
    theta = hist[-1][1] @ hist[-1][1] / (hist[-1][0] @ hist[-1][1])
    M = np.diag(rng.uniform(1, 2, n))
    h0 = (lambda v: np.linalg.solve(M, v)) if use_h0 else None
    H0 = np.linalg.inv(M) if use_h0 else np.eye(n) / theta
    v = rng.standard_normal(n)
    res = O.two_loop(S, Y, ys, theta, nc, ptr, v, a=-1.0, h0=h0)
    Hd = dense_bfgs_inverse([h[0] for h in hist], [h[1] for h in hist], H0)
    assert np.allclose(res, -Hd @ v, rtol=1e-10, atol=1e-12)


def test_lbfgs_quadratic_terminates(rng):
    n = 10
    A = rng.standard_normal((n, n))
    A = A @ A.T + np.eye(n)
    b = rng.standard_normal(n)
    x, st = O.lbfgs(n, lambda x: A @ x - b, np.zeros(n), h0=lambda v: np.linalg.solve(A, v), epsilon=1e-12)
    # exact H0 = A^-1: the first unit step lands on the solution
    assert st["iterations"] == 1 and st["r_evals"] == 2
    assert np.allclose(x, np.linalg.solve(A, b))
    x, st = O.lbfgs(n, lambda x: A @ x - b, np.zeros(n), m=20, epsilon=1e-10, max_iterations=500)
    assert np.allclose(A @ x, b, atol=1e-9)


def test_projection_line_search_trajectory():
    # scalar r(x) = x^3 - 8 from x0 = 0 with H0 = I (theta = 1 on the first
    # step: d = -r = 8).  Trials: x = a d, accepted when |r(x) d| <= |r(0) d|;
    # otherwise a *= 1.5 if sign(r d) == sign(r0 d), else 0.5
    r = lambda x: x ** 3 - 8.0
    d, r0 = 8.0, -8.0
    a, exp_trials = 1.0, 0
    while True:
        exp_trials += 1
        rr = r(a * d)
        if abs(rr * d) <= abs(r0 * d):
            break
        a *= 1.5 if rr * d * r0 * d > 0 else 0.5
    x, st = O.lbfgs(1, lambda x: r(x), np.zeros(1), epsilon=0.0, max_iterations=1, trace_cap=4)
    # max_iterations = 1 -> stop right after the first line search
    assert st["ls_trials"] == exp_trials
    assert np.isclose(st["trace"][0, 1], a)
    assert np.isclose(x[0], a * d)


def test_line_search_accepts_last_trial_silently():
    # |r(x)| grows away from the origin faster than |r(x0)|: every trial of
    # the first line search fails the projection test
    r = lambda x: np.sign(x) * 10.0 * (1.0 + np.abs(x))
    x, st = O.lbfgs(1, r, np.array([-1.0]), max_linesearch=3, max_iterations=1, epsilon=0.0)
    assert st["ls_trials"] == 3 and st["iterations"] == 1


def test_early_exit_uses_l2_norm_of_x():
    # ||r||_inf <= eps_rel * ||x||_2 passes before the loop (LBFGS.h:118)
    x0 = np.array([3.0, 4.0])
    x, st = O.lbfgs(2, lambda x: np.array([4.9e-6, 0.0]), x0, epsilon=1e-9, epsilon_rel=1e-6)
    assert st["return"] == 1 and st["r_evals"] == 1
    x, st = O.lbfgs(2, lambda x: np.array([5.1e-6, 0.0]), x0, epsilon=1e-9, epsilon_rel=1e-6,
                    max_iterations=1)
    assert st["r_evals"] > 1


def test_model_solve_converges_small_load():
    m = E.Mesh.workload(0, [4, 2, 2, 0.004, 0.002, 0.002, 8, 0, 50.0])
    om = O.Model(m)
    u, st = om.solve(epsilon=1e-10, max_iterations=2000)
    assert st["rinf"] <= 1e-10
    assert np.abs(om.residual(u)).max() <= 1e-10
    assert st["precond_failed"] == 0


def test_all_fixed_mesh_falls_back_without_factor():
    # K with no entries -> the reference logs and uses H0 = I/theta (LBFGS.h:71-80)
    m = E.Mesh.workload(0, [2, 1, 1, 0.002, 0.001, 0.001, 8, 0, 50.0])
    m.fixed[:] = 1
    om = O.Model(m)
    assert om.ndof == 0
    assert not om.factor()
