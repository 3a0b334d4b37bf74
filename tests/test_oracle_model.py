"""Pins the oracle's residual, pressure and approximate Jacobian by
invariants and closed forms.

Reference: src/fe/ElementConstraint.cpp:139-205, src/fe/ExternalConstraint.cpp:
249-959, src/fe/Constraint.cpp:61-126.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import efem_amd as E
import oracle as O

KINDS = [(8, 0), (8, 1), (4, 0), (10, 0), (10, 1), (20, 0), (20, 1)]


def box(kind, reduced=0, law=0, nx=3, ny=2, nz=2, p=500.0, jitter=0.0):
    # custom box {nx, ny, nz, lx, ly, lz, kind, law, pressure, jitter, seed, reduced}
    return E.Mesh.workload(0, [nx, ny, nz, 0.003, 0.002, 0.002, kind, law, p, jitter, 5, reduced])


def free_box(kind, reduced=0, law=0):
    m = box(kind, reduced, law)
    return E.Mesh(m.X, m.conn, m.law, m.params, np.zeros(m.n_nodes, np.uint8), None, 0.0, m.reduced)


@pytest.mark.parametrize("kind,reduced", KINDS)
def test_rigid_motion_gives_zero_internal_force(kind, reduced, rng):
    m = free_box(kind, reduced, law=1)
    om = O.Model(m)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    u = (m.X @ Q.T - m.X + rng.standard_normal(3)).ravel()
    f = om.internal_force(u)
    assert np.abs(f).max() < 1e-9 * 1e5 * 1e-6  # moduli x area scale


@pytest.mark.parametrize("kind,reduced", KINDS)
def test_patch_test_linear_field(kind, reduced, rng):
    # homogeneous deformation: uniform P, interior nodal forces vanish and
    # the boundary forces sum to zero
    m = free_box(kind, reduced, law=0)
    om = O.Model(m)
    A = 0.05 * rng.standard_normal((3, 3))
    u = (m.X @ A.T).ravel()
    f = om.internal_force(u).reshape(-1, 3)
    X = m.X
    lo, hi = X.min(0), X.max(0)
    interior = np.all((X > lo + 1e-9) & (X < hi - 1e-9), axis=1)
    scale = np.abs(f).max()
    assert scale > 0
    if interior.any():
        assert np.abs(f[interior]).max() < 1e-10 * scale
    assert np.abs(f.sum(0)).max() < 1e-10 * scale
    # and each boundary node carries P . (its share of the reference normal area):
    # total force on the x = max face equals P @ (face area * e_x)
    P = O.piola(0, m.params, np.eye(3) + A)
    xmax = np.isclose(X[:, 0], hi[0])
    area = (hi[1] - lo[1]) * (hi[2] - lo[2])
    assert np.allclose(f[xmax].sum(0), P @ np.array([area, 0, 0]), rtol=1e-9, atol=1e-12 * scale)


@pytest.mark.parametrize("kind,reduced", [(8, 0), (4, 0), (10, 0), (20, 1)])
def test_pressure_undeformed_closed_form(kind, reduced):
    m = box(kind, reduced, p=700.0)
    om = O.Model(m)
    fe = -om.external_force(np.zeros(m.ndof)).reshape(-1, 3)  # +f_ext on free DOFs
    fid = E.free_ids(m.fixed)
    # all pressure pushes along -z on the z = max face, total p * A (minus
    # what lands on fixed nodes); every face node's share is p * int N_a
    assert np.abs(fe[:, :2]).max() < 1e-12
    lo, hi = m.X.min(0), m.X.max(0)
    A_face = (hi[0] - lo[0]) * (hi[1] - lo[1])
    top = np.isclose(m.X[:, 2], hi[2])
    on_fixed = top & (m.fixed == 1)
    if kind in (8, 4):
        # linear faces: p * A_f / n_corners per corner per face
        tot_fixed = 0.0
        for f in m.faces:
            share = 700.0 * (A_face / len(m.faces)) / len(f)
            tot_fixed += share * m.fixed[f].sum()
        assert abs(-fe[:, 2].sum() - (700.0 * A_face - tot_fixed)) < 1e-12
    else:
        # quadratic faces: every node's exact share, see face_share_closed_form
        want = face_share_closed_form(m, 700.0)
        free = m.fixed == 0
        assert np.allclose(fe[fid[free], 2], want[free, 2], rtol=0, atol=1e-12 * 700.0 * A_face)
    assert on_fixed.any()
    # nodes off the face get N_a(xi_q) = 0 up to rounding of 1 - xi - eta - zeta
    assert np.abs(fe[fid[~top & (m.fixed == 0)], 2]).max() < 1e-15 * 700.0 * A_face


def face_share_closed_form(m, p):
    """+f_ext per node of a follower pressure p on flat, straight-edged
    faces at u = 0, from the exact integrals of the face shape functions:
    f_a = p A_f int N_a dA / A_f along the inward normal, where
    int N_a / A is 1/3 per corner of a 3-node triangle, 1/4 per corner of a
    bilinear quad, 0 per corner and 1/3 per mid-edge node of a 6-node
    triangle, and -1/12 per corner and 1/3 per mid-edge node of an 8-node
    serendipity quad.  Corners are the first 3 (4) face nodes.  The inward
    normal is -z for the top faces of the box meshes."""
    fnn = m.faces.shape[1]
    share = {3: [1 / 3] * 3, 4: [1 / 4] * 4, 6: [0.0] * 3 + [1 / 3] * 3,
             8: [-1 / 12] * 4 + [1 / 3] * 4}[fnn]
    nc = 3 if fnn in (3, 6) else 4
    out = np.zeros((m.n_nodes, 3))
    for f in m.faces:
        P = m.X[f[:nc]]
        if nc == 3:
            n = 0.5 * np.cross(P[1] - P[0], P[2] - P[0])
        else:
            n = 0.5 * np.cross(P[2] - P[0], P[3] - P[1])
        area = np.linalg.norm(n)
        for a, s in zip(f, share):
            out[a] += p * area * s * np.array([0.0, 0.0, -1.0])
    return out


@pytest.mark.parametrize("kind,reduced", [(8, 0), (4, 0), (10, 0), (20, 1), (20, 0)])
def test_pressure_shares_under_homogeneous_deformation(kind, reduced, rng):
    # u = (F - I) X: Nanson gives every face point the same cof(F) A0, so
    # each node's follower force is cof(F) times its undeformed closed-form
    # share (ExternalConstraint.cpp:829-836) -- for all four face kinds,
    # with the face rule's points and weights entering only through the
    # exact integrals of N_a
    m = box(kind, reduced, p=700.0)
    m = E.Mesh(m.X, m.conn, m.law, m.params, np.zeros(m.n_nodes, np.uint8), m.faces, m.pressure, m.reduced)
    om = O.Model(m)
    F = np.eye(3) + 0.15 * rng.standard_normal((3, 3))
    assert np.linalg.det(F) > 0.3
    u = (m.X @ (F - np.eye(3)).T).ravel()
    got = -om.external_force(u).reshape(-1, 3)  # + f_ext
    cof = np.linalg.det(F) * np.linalg.inv(F).T
    want = face_share_closed_form(m, 700.0) @ cof.T
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-12 * scale * 10


def test_pressure_on_closed_surface_is_self_equilibrated(rng):
    # follower pressure over the whole boundary of a free body: zero resultant
    # and zero moment at any deformation (Nanson: sum cof(F) A0 = 0)
    m0 = box(8, 0, nx=2, ny=2, nz=2)
    X = m0.X
    faces = []
    lo, hi = X.min(0), X.max(0)
    for e in m0.conn:
        for fl in ((0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7)):
            fn = e[list(fl)]
            P = X[fn]
            for ax in range(3):
                for v in (lo[ax], hi[ax]):
                    if np.allclose(P[:, ax], v):
                        faces.append(fn)
    m = E.Mesh(X, m0.conn, 0, m0.params, np.zeros(len(X), np.uint8), np.array(faces), 1000.0)
    om = O.Model(m)
    A = 0.1 * rng.standard_normal((3, 3))
    u = (X @ A.T).ravel()
    f = om.external_force(u).reshape(-1, 3)
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).max() * len(f)


def test_pressure_is_a_follower_load(rng):
    # rotating the body rotates the pressure forces (ExternalConstraint.cpp:829-836)
    m = box(10, 0, p=900.0)
    m = E.Mesh(m.X, m.conn, m.law, m.params, np.zeros(m.n_nodes, np.uint8), m.faces, m.pressure)
    om = O.Model(m)
    f0 = om.external_force(np.zeros(m.ndof)).reshape(-1, 3)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    u = (m.X @ Q.T - m.X).ravel()
    f1 = om.external_force(u).reshape(-1, 3)
    assert np.allclose(f1, f0 @ Q.T, atol=1e-12 * np.abs(f0).max())


@pytest.mark.parametrize("kind,reduced", KINDS)
def test_approximate_jacobian_structure(kind, reduced):
    m = box(kind, reduced, law=2)
    om = O.Model(m)
    K = om.K(scalar=False).toarray()
    L = om.K(scalar=True).toarray()
    assert np.allclose(K, K.T, rtol=0, atol=0)
    assert np.allclose(K, np.kron(L, np.eye(3)), rtol=0, atol=0)
    # free-free Laplacian: a constant field is in the kernel before BCs
    mf = free_box(kind, reduced, law=2)
    Lf = O.Model(mf).K(scalar=True).toarray()
    assert np.abs(Lf @ np.ones(len(Lf))).max() < 1e-10 * np.abs(Lf).max()
    # SPD after eliminating the fixed nodes -- except for the one-point
    # reduced rules of quadratic tets and linear hexes, whose Laplacians have
    # hourglass (zero-energy) modes
    ev = np.linalg.eigvalsh(L)
    if (kind, reduced) in ((8, 1), (10, 1)):
        assert ev.min() < 1e-9 * ev.max()
    else:
        assert ev.min() > 0


def test_approximate_jacobian_coefficient_and_weights():
    # one free unit-cube C3D8 element: L = V c sum_q w_q gradN gradN^T with
    # c = 2.94001 C10 + 2/D1 (NeoHookean.cpp:41); check trace against the
    # closed form of the 2x2x2-integrated Laplacian of the unit cube
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    m = E.Mesh(X, np.arange(8)[None], 0, [1.0, 0, 0, 2.0], np.zeros(8, np.uint8))
    L = O.Model(m).K(scalar=True).toarray()
    c = 2.94001 * 1.0 + 2.0 / 2.0
    # integrated Laplacian stiffness of the unit cube, diagonal = 1/3
    assert np.allclose(np.diag(L), c / 3.0, rtol=1e-14)


def test_cholesky_preconditioner_solves_K(rng):
    m = box(4, 0, nx=4, ny=3, nz=3, jitter=0.1)
    om = O.Model(m)
    K = om.K(scalar=False)
    b = rng.standard_normal(m.ndof)
    x = spla.spsolve(K.tocsc(), b)
    for scalar in (False, True):
        assert om.factor(scalar=scalar)
        assert np.linalg.norm(om.apply_h0(b) - x) <= 1e-12 * np.linalg.norm(x)


def test_free_dof_numbering():
    m = box(8)
    om = O.Model(m)
    fid = E.free_ids(m.fixed)
    for n in range(m.n_nodes):
        assert O.lib().or_free_id(om._h, n) == fid[n]
