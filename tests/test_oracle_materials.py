"""Pins the oracle's constitutive laws against the reference's own sources
(golden vectors from oracle/_ref, tests/golden/material_ref.npz) and against
the mathematics (P(I) = 0, P = dW/dF, objectivity).

Reference: src/material/{NeoHookean,MooneyRivlin,Yeoh}.cpp.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLD = Path(__file__).parent / "golden" / "material_ref.npz"


def test_oracle_piola_matches_reference_sources():
    g = np.load(GOLD)
    F = g["F"]
    nF = len(F)
    for k, (law, p, Pref) in enumerate(zip(g["law"], g["params"], g["P"])):
        P = O.piola(int(law), p, F[k % nF])
        # near F = I the stress is a difference of modulus-sized terms, so
        # rounding is relative to the moduli, not to |P|
        moduli = np.abs(p[:3]).sum()
        assert np.abs(P - Pref).max() <= 1e-12 * np.abs(Pref).max() + 1e-14 * moduli, (law, p, k)


def test_elasticity_coeff_matches_reference_sources():
    g = np.load(GOLD)
    for row in g["coeff"]:
        law, p, c = int(row[0]), row[1:5], row[5]
        assert O.lib().or_elasticity_coeff(law, O._p(np.ascontiguousarray(p))) == pytest.approx(c, rel=1e-15)


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
def test_oracle_piola_matches_reference_live(rng):
    for law, p in ((0, [114000, 0, 0, 1e-6]), (1, [80000, 20000, 0, 1e-6]), (2, [5e5, -1e4, 1e3, 1e-7])):
        for _ in range(20):
            F = np.eye(3) + 0.4 * rng.standard_normal((3, 3))
            if np.linalg.det(F) < 0.1:
                continue
            Pr = O.ref_piola(law, p, F)
            tol = 1e-12 * np.abs(Pr).max() + 1e-14 * np.abs(p[:3]).sum()
            assert np.abs(O.piola(law, p, F) - Pr).max() <= tol


def _energy(law, p, F):
    # strain energies whose derivatives are the reference's P (no counterpart
    # in the reference, which never evaluates an energy)
    J = np.linalg.det(F)
    C = F.T @ F
    I1 = np.trace(C)
    I1b = I1 * J ** (-2 / 3)
    W = (J - 1) ** 2 / p[3]
    if law == 2:
        t = I1b - 3
        return W + p[0] * t + p[1] * t ** 2 + p[2] * t ** 3
    W += p[0] * (I1b - 3)
    if law == 1:
        I2 = 0.5 * (I1 ** 2 - np.trace(C @ C))
        W += p[1] * (I2 * J ** (-4 / 3) - 3)
    return W


@pytest.mark.parametrize("law,p", [(0, [114000, 0, 0, 1e-3]), (1, [80000, 20000, 0, 1e-3]),
                                   (2, [5e5, -1e4, 1e3, 1e-3])])
def test_piola_is_energy_gradient(law, p, rng):
    p = np.array(p, float)
    h = 1e-7
    for _ in range(5):
        F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        P = O.piola(law, p, F)
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                G[i, j] = (_energy(law, p, F + E) - _energy(law, p, F - E)) / (2 * h)
        assert np.abs(P - G).max() <= 2e-6 * np.abs(P).max()


@pytest.mark.parametrize("law", [0, 1, 2])
def test_stress_free_reference_and_objectivity(law, rng):
    p = np.array([1e5, 2e4 if law == 1 else (-1e3 if law == 2 else 0), 1e2 if law == 2 else 0, 1e-5])
    assert np.abs(O.piola(law, p, np.eye(3))).max() < 1e-9
    for _ in range(5):
        F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        assert np.allclose(O.piola(law, p, Q @ F), Q @ O.piola(law, p, F), rtol=1e-10,
                           atol=1e-10 * np.abs(O.piola(law, p, F)).max())


def test_uniaxial_incompressible_limit_closed_form():
    # Neo-Hookean, isochoric stretch F = diag(l, l^-1/2, l^-1/2) (J = 1):
    # P11 - P22 * (F22/F11) = 2 C10 (l - l^-2)  (Cauchy sigma11 - sigma22 = 2C10(l^2 - 1/l))
    C10 = 1e5
    for lam in (0.7, 1.2, 2.0):
        F = np.diag([lam, lam ** -0.5, lam ** -0.5])
        P = O.piola(0, [C10, 0, 0, 1e-6], F)
        s11, s22 = P[0, 0] * lam, P[1, 1] * lam ** -0.5
        assert abs((s11 - s22) - 2 * C10 * (lam ** 2 - 1 / lam)) < 1e-6 * C10
