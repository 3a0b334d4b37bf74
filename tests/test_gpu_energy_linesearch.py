"""GPU: the energy line search (EFEM_LS_ENERGY, an addition: the reference
runs only the projection rule of src/simulation/nlsolver/LineSearchFEA.h:
37-68) against the oracle's energy rule, on conservative loads.

  * a cantilever sagging under its own weight (gravity, consistent mass):
    the first iterations follow the oracle's energy-rule path (same trial
    counts and steps), and the solve reaches the projection rule's
    equilibrium within 1e-6;
  * with a dead nodal force as well, in mixed precision;
  * refused (EFEM_ERR_ARG) when the problem has pressure faces.
"""
import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import rel

pytestmark = pytest.mark.gpu


def self_weight(nx=20, fnodes=None, fvec=None):
    m0 = E.Mesh.workload(1, [nx, 4, 4, 0, 300.0])
    return E.Mesh(m0.X, m0.conn, m0.law, m0.params, m0.fixed, None, 0.0, m0.reduced, gravity=6,
                  consistent_inertia=1, density=[1050.0], fnodes=fnodes, fvec=fvec)


def test_energy_rule_follows_the_oracle_and_reaches_the_equilibrium():
    m = self_weight()
    om = O.Model(m)
    print("energy rule: self-weight cantilever ndof", m.ndof, flush=True)
    _, st_or = om.solve(epsilon=0.0, max_iterations=10, trace_cap=10, line_search=1)
    with E.Context(m, E.FP64) as ctx:
        _, st = ctx.solve(epsilon=0.0, max_iterations=10, line_search=E.LS_ENERGY)
        tr = ctx.trace()
        assert st["status"] == 0 and st["iterations"] == 10
        assert np.array_equal(tr[:, 1:], st_or["trace"][:, 1:])
        assert np.allclose(tr[:, 0], st_or["trace"][:, 0], rtol=1e-6)
        u_p, st_p = ctx.solve(epsilon=1e-9, max_iterations=5000)
        u_e, st_e = ctx.solve(epsilon=1e-9, max_iterations=5000, line_search=E.LS_ENERGY)
    print("projection", st_p["iterations"], "energy", st_e["iterations"], flush=True)
    assert st_p["rinf"] <= 1e-9 and st_e["rinf"] <= 1e-9
    assert rel(u_e, u_p) <= 1e-6
    # the potential went down along the accepted steps
    assert om.potential(u_e) < om.potential(np.zeros(m.ndof))


def test_energy_rule_rejects_what_the_projection_rule_accepts():
    # lns_fea = 1e6 makes the projection test pass at every trial and
    # H0 = I/theta starts from d = -r: step 1 raises the potential, so the
    # energy rule halves it (12 trials, step 2^-11 on the oracle) where the
    # projection rule takes it; the device follows the oracle trial by trial
    m = self_weight(10)
    om = O.Model(m)
    kw = dict(epsilon=0.0, max_iterations=8, use_precond=False, lns_fea=1e6)
    _, st_or = om.solve(trace_cap=8, line_search=1, **kw)
    _, st_pr = om.solve(trace_cap=8, line_search=0, **kw)
    assert st_or["trace"][0, 2] > 1 and st_pr["trace"][0, 2] == 1
    with E.Context(m, E.FP64) as ctx:
        _, st = ctx.solve(line_search=E.LS_ENERGY, **kw)
        tr = ctx.trace()
    assert np.array_equal(tr[:, 1:], st_or["trace"][:, 1:])
    assert np.allclose(tr[:, 0], st_or["trace"][:, 0], rtol=1e-6)


def test_energy_rule_mixed_precision_with_a_dead_nodal_force():
    m0 = self_weight(12)
    tip = np.flatnonzero(np.isclose(m0.X[:, 0], m0.X[:, 0].max()))
    m = self_weight(12, fnodes=tip, fvec=np.array([0.0, 0.002, -0.001]))
    with E.Context(m, E.FP64) as ctx:
        u64, st64 = ctx.solve(epsilon=1e-10, max_iterations=5000)
    with E.Context(m, E.MIXED) as ctx:
        u, st = ctx.solve(epsilon=1e-8, max_iterations=5000, line_search=E.LS_ENERGY, fp64_switch=1e-3)
    print("mixed energy rule", st["iterations"], st["rinf"], flush=True)
    assert st["status"] == 0 and st["rinf"] <= 1e-8
    assert rel(u, u64) <= 1e-5


def test_energy_rule_refused_with_pressure_faces():
    m = E.Mesh.workload(1, [6, 2, 2, 0, 100.0])
    with E.Context(m, E.FP64) as ctx:
        with pytest.raises(E.EfemError, match="conservative"):
            ctx.solve(epsilon=1e-6, max_iterations=5, line_search=E.LS_ENERGY)
