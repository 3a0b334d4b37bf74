"""Material tangent and strain energy of the three laws (host entry point
efem_material_point, no GPU).

The tangent dP = (dP/dF)[dF] comes from forward-mode dual numbers through
the same stress code the element kernels run.  It is pinned by
  * central differences of P (a dropped or mis-signed term shows at O(1));
  * the small-strain limit at F = I, a closed form:
      dP = mu (dF + dF^T - 2/3 tr dF I) + kappa tr dF I,
      mu = 2 (C10 + C01), kappa = 2 / D1;
  * major symmetry dF2 : A dF1 = dF1 : A dF2 (hyperelasticity);
  * dW = P : dF by central differences, W(I) = 0, P(I) = 0;
  * objectivity W(RF) = W(F), P(RF) = R P(F).
"""
import numpy as np
import pytest

import efem_amd as E

LAWS = [(E.NEOHOOKEAN, [114000.0, 0.0, 0.0, 1e-6]),
        (E.MOONEY_RIVLIN, [80000.0, 20000.0, 0.0, 1e-6]),
        (E.YEOH, [5e5, -1e4, 1e3, 1e-7])]


def rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@pytest.mark.parametrize("law,p", LAWS)
@pytest.mark.parametrize("amp", [1e-3, 0.1, 0.3])
def test_tangent_matches_central_differences(law, p, amp, rng):
    for _ in range(5):
        F = np.eye(3) + amp * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.5:
            continue
        dF = rng.standard_normal((3, 3))
        _, _, dP = E.material_point(law, p, F, dF)
        h = 1e-6
        Pp, _, _ = E.material_point(law, p, F + h * dF)
        Pm, _, _ = E.material_point(law, p, F - h * dF)
        fd = (Pp - Pm) / (2 * h)
        assert np.abs(dP - fd).max() <= 1e-6 * np.abs(dP).max()


@pytest.mark.parametrize("law,p", LAWS)
def test_small_strain_limit(law, p, rng):
    mu = 2.0 * (p[0] + (p[1] if law == E.MOONEY_RIVLIN else 0.0))
    kappa = 2.0 / p[3]
    for _ in range(4):
        dF = rng.standard_normal((3, 3))
        _, _, dP = E.material_point(law, p, np.eye(3), dF)
        tr = np.trace(dF)
        want = mu * (dF + dF.T - 2.0 / 3.0 * tr * np.eye(3)) + kappa * tr * np.eye(3)
        assert np.abs(dP - want).max() <= 1e-9 * np.abs(want).max()


@pytest.mark.parametrize("law,p", LAWS)
def test_major_symmetry(law, p, rng):
    for _ in range(5):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.5:
            continue
        a, b = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
        _, _, Aa = E.material_point(law, p, F, a)
        _, _, Ab = E.material_point(law, p, F, b)
        lhs, rhs = (b * Aa).sum(), (a * Ab).sum()
        assert abs(lhs - rhs) <= 1e-9 * (abs(lhs) + abs(rhs))


@pytest.mark.parametrize("law,p", LAWS)
def test_energy_gradient_is_piola(law, p, rng):
    P0, W0, _ = E.material_point(law, p, np.eye(3))
    assert W0 == 0.0 and np.abs(P0).max() == 0.0
    for _ in range(5):
        F = np.eye(3) + 0.15 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.6:
            continue
        dF = rng.standard_normal((3, 3))
        P, W, _ = E.material_point(law, p, F)
        h = 1e-7
        _, Wp, _ = E.material_point(law, p, F + h * dF)
        _, Wm, _ = E.material_point(law, p, F - h * dF)
        fd = (Wp - Wm) / (2 * h)
        want = (P * dF).sum()
        assert abs(fd - want) <= 1e-6 * max(abs(want), np.abs(P).max())


@pytest.mark.parametrize("law,p", LAWS)
def test_objectivity(law, p, rng):
    for _ in range(5):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        R = rot(rng)
        P, W, _ = E.material_point(law, p, F)
        PR, WR, _ = E.material_point(law, p, R @ F)
        assert abs(WR - W) <= 1e-9 * max(abs(W), 1.0)
        assert np.abs(PR - R @ P).max() <= 1e-9 * np.abs(P).max()


def test_bad_arguments():
    with pytest.raises(E.EfemError):
        E.material_point(7, [1.0, 0, 0, 1.0], np.eye(3))


@pytest.mark.parametrize("law,p", LAWS)
def test_inverted_element_stress_matches_oracle(law, p, rng):
    # J < 0 (an element turned inside out by a large trial step): the
    # reference's invariants use (J^2)^(1/3) = |J|^(2/3) and stay finite
    # (NeoHookean.cpp:9-10), so must the kernels' stress
    import oracle as O
    for _ in range(10):
        F = rng.standard_normal((3, 3))
        if np.linalg.det(F) > -0.2:
            F[:, 0] *= -1
        if abs(np.linalg.det(F)) < 0.2:
            continue
        P, _, _ = E.material_point(law, p, F)
        Po = O.piola(law, np.asarray(p), F)
        assert np.all(np.isfinite(P))
        scale = np.abs(Po).max() + 1e-14 * np.abs(p[:3]).sum()
        assert np.abs(P - Po).max() <= 1e-12 * scale
