"""Gravity (the reference's InertiaConstraint): the constant force M g over
the free DOFs, M consistent or lumped.  The host model's force and the
oracle's independent restatement are pinned by closed-form mass matrices
(linear tetrahedron, trilinear hexahedron), the lumped mass floor and
partition of unity, then compared on random meshes."""
import json

import numpy as np
import pytest

import efem_amd as E
import oracle as O

RHO = 1050.0
G = 9.8


def tet(fixed=(), gravity=6, consistent=1, rho=RHO):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    fx = np.zeros(4, np.uint8)
    fx[list(fixed)] = 1
    return E.Mesh(X, [[0, 1, 2, 3]], E.NEOHOOKEAN, [1e5, 0, 0, 1e-6], fx, gravity=gravity,
                  consistent_inertia=consistent, density=[rho])


def cube(fixed_x0=True, consistent=1):
    X = np.array([[i, j, k] for k in (0, 1) for j in (0, 1) for i in (0, 1)], float)
    # C3D8 node order: bottom face counter-clockwise, then top face
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    X = np.array(order, float)
    fx = (X[:, 0] == 0).astype(np.uint8) if fixed_x0 else np.zeros(8, np.uint8)
    return E.Mesh(X, [list(range(8))], E.NEOHOOKEAN, [1e5, 0, 0, 1e-6], fx, gravity=6,
                  consistent_inertia=consistent, density=[RHO])


def oracle_force(m):
    # the oracle's residual at u = 0 without minus with gravity: the internal
    # force (rounding at F = I) and any nodal forces cancel bit for bit
    u0 = np.zeros(3 * m.n_free)
    g = m.gravity
    with_g = O.Model(m).residual(u0)
    m.gravity = 0
    without = O.Model(m).residual(u0)
    m.gravity = g
    return without - with_g


def per_node(f):
    return f.reshape(-1, 3)


@pytest.mark.parametrize("consistent", [1, 0])
def test_single_tet_closed_forms(consistent):
    V = 1.0 / 6.0
    # all nodes free: every row of either mass matrix sums to rho V / 4
    m = tet(consistent=consistent)
    for f in (E.host_constant_force(m), oracle_force(m)):
        assert np.allclose(per_node(f), [[0, 0, -RHO * V * G / 4]] * 4, rtol=1e-14, atol=0)
    # node 3 fixed: the reference integrates C3D4 with its one-point rule
    # (N_a = 1/4 at the centroid), so every consistent entry is rho V / 16
    # and a free node keeps 3 / 16 of rho V (an exact rule would give the
    # textbook rho V (1 + delta_ab) / 20); lumped stays rho V / 4
    m = tet(fixed=(3,), consistent=consistent)
    share = 3.0 / 16.0 if consistent else 1.0 / 4.0
    for f in (E.host_constant_force(m), oracle_force(m)):
        assert np.allclose(per_node(f), [[0, 0, -RHO * V * G * share]] * 3, rtol=1e-14, atol=0)


def test_cube_consistent_mass_tensor_product():
    # trilinear hexahedron: M_ab = rho V / 216 * prod_d (2 if a, b share
    # coordinate d else 1) (the 1-D mass h/6 [2 1; 1 2] per direction); the
    # x = 1 face is free: 8 + 4 + 4 + 2 = 18 / 216 of rho V per free node
    m = cube()
    for f in (E.host_constant_force(m), oracle_force(m)):
        assert np.allclose(per_node(f), [[0, 0, -RHO * G * 18.0 / 216.0]] * 4, rtol=1e-13, atol=0)
    m = cube(consistent=0)
    for f in (E.host_constant_force(m), oracle_force(m)):
        assert np.allclose(per_node(f), [[0, 0, -RHO * G / 8.0]] * 4, rtol=1e-14, atol=0)


def test_axes_and_lumped_floor():
    V = 1.0 / 6.0
    for gv, axis, sign in ((1, 0, 1), (2, 0, -1), (3, 1, 1), (4, 1, -1), (5, 2, 1), (6, 2, -1)):
        f = per_node(E.host_constant_force(tet(gravity=gv, consistent=0)))
        want = np.zeros(3)
        want[axis] = sign * RHO * V * G / 4
        assert np.allclose(f, [want] * 4, rtol=1e-14, atol=0)
    # ElementConstraint.cpp:108-111: the element mass is floored at 4e-12
    f = per_node(E.host_constant_force(tet(consistent=0, rho=1e-15)))
    assert np.allclose(f[:, 2], -4e-12 / 4 * G, rtol=1e-14)
    assert np.allclose(per_node(oracle_force(tet(consistent=0, rho=1e-15)))[:, 2], -4e-12 / 4 * G, rtol=1e-14)
    # gravity 0, or no density: no force
    assert not E.host_constant_force(tet(gravity=0)).any()
    m = tet()
    m.density = None
    assert not E.host_constant_force(m).any()


@pytest.mark.parametrize("kind,law,reduced", [(4, 0, 0), (10, 0, 0), (8, 1, 0), (20, 2, 0), (20, 2, 1)])
@pytest.mark.parametrize("consistent", [1, 0])
def test_random_mesh_matches_oracle(kind, law, reduced, consistent, rng):
    m = E.Mesh.workload(0, [4, 3, 2, 0.004, 0.003, 0.002, kind, law, 1.0, 0.1 if kind in (4, 10) else 0.0,
                            5, reduced])
    m = E.Mesh(m.X, m.conn, m.law, m.params, m.fixed, None, 0.0, m.reduced, gravity=3,
               consistent_inertia=consistent, density=rng.uniform(900, 1200, m.n_elems))
    m.params = np.tile(np.atleast_2d(m.params), (m.n_elems, 1))
    m.params_per_elem = 1
    f, fo = E.host_constant_force(m), oracle_force(m)
    assert np.linalg.norm(f - fo) <= 1e-13 * np.linalg.norm(fo)
    # partition of unity: with every node free the total is rho V g summed
    # (the fixed face here keeps part of the mass: the total is below it)
    total = sum(m.density[e] * O.Model(m).volume(e) for e in range(m.n_elems)) * G
    assert 0 < f.reshape(-1, 3)[:, 1].sum() <= total * (1 + 1e-12)


def test_gravity_deck(tmp_path):
    d = tmp_path / "g"
    d.mkdir()
    (d / "mesh.txt").write_text("#node\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n#element\n1 2 3 4 5\n")
    cfg = {"model": [{"type": "soft", "mesh_file": "mesh.txt"}], "constraint": [],
           "material": [{"id": 5, "type": "neo_hookean", "C10": 1e5, "D1": 1e-6, "density": 1050}],
           "analysis": {"gravity": "-z", "consistent_inertia": True}}
    (d / "run.json").write_text(json.dumps(cfg))
    m, _ = E.Mesh.from_input(d / "run.json")
    assert m.gravity == 6 and m.consistent_inertia == 1 and np.allclose(m.density, [1050])
    f = per_node(E.host_constant_force(m))
    assert np.allclose(f, [[0, 0, -1050 / 6 * 9.8 / 4]] * 4, rtol=1e-14)
    cfg["analysis"] = {"gravity": "non"}
    (d / "run.json").write_text(json.dumps(cfg))
    m, _ = E.Mesh.from_input(d / "run.json")
    assert m.gravity == 0
