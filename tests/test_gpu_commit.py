"""GPU: committed configurations (Constraint::updateDx, src/fe/Constraint.cpp:
103-113) and the run driver with its VTK frames (Simulator.cpp:6-40,
Model::saveVTK).

The reference solves for the increment dx from 0 with the residual taken at
node->dx + dx and stops on ||r||_inf <= epsilon_rel ||dx||_inf; the device
solve iterates on the total displacement from the committed one.  The
oracle restates the reference's form literally -- its L-BFGS on dx with the
residual at u_c + dx -- and the device must follow the same iterates and
stop by the same rule."""
import shutil
from pathlib import Path

import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import rel
from test_vtk import read_legacy_vtk, scatter

pytestmark = pytest.mark.gpu

SAMPLE = Path(__file__).resolve().parent / "golden" / "sample_beam"


def test_commit_state():
    m = E.Mesh.workload(1, [10, 2, 2, 0, 100.0])
    with E.Context(m) as c:
        assert not c.committed().any()
        u = np.linspace(-1e-4, 1e-4, c.ndof)
        c.commit(u)
        assert np.array_equal(c.committed(), u)
        c.commit(None)
        assert not c.committed().any()
        with pytest.raises(ValueError):
            c.commit(np.zeros(c.ndof + 3))
    with E.Batch([m, m]) as b:
        with pytest.raises(E.EfemError, match="batch"):
            b.commit(np.zeros(b.ndof))


def test_increment_solve_matches_reference_form():
    m = E.Mesh.workload(1, [20, 4, 4, 0, 300.0])
    n = m.ndof
    print(f"cantilever 20x4x4 C3D8, {n} DOF", flush=True)
    # the stop is decided by the relative criterion alone
    kw = dict(epsilon=1e-14, epsilon_rel=1e-7, max_iterations=3000)
    with E.Context(m, E.FP64) as c:
        u1, s1 = c.solve(epsilon=1e-14, epsilon_rel=0.0, max_iterations=40)
        assert s1["iterations"] == 40
        # without a commit the criterion is measured from u = 0
        u_nc, s_nc = c.solve(u1, **kw)
        assert s_nc["xinf"] == np.abs(u_nc).max()
        c.commit(u1)
        u2, s2 = c.solve(u1, **kw)
        tr = c.trace()
    print("no commit", s_nc["iterations"], "commit", s2["iterations"], flush=True)
    assert s2["status"] == 0
    assert s2["xinf"] == np.abs(u2 - u1).max()
    assert s2["rinf"] <= kw["epsilon_rel"] * s2["xinf"]
    # the reference's form on the oracle: dx from 0, residual at u1 + dx
    om = O.Model(m)
    assert om.factor()
    dx, so = O.lbfgs(n, lambda d: om.residual(u1 + d), np.zeros(n), h0=om.apply_h0, trace_cap=10, **kw)
    print("oracle", so["iterations"], so["rinf"], flush=True)
    assert so["rinf"] <= kw["epsilon_rel"] * np.abs(dx).max()
    # the same first iterates; a comparable count; the same equilibrium
    assert np.allclose(tr[:10, 0], so["trace"][:10, 0], rtol=1e-6)
    assert abs(s2["iterations"] - so["iterations"]) <= max(5, 0.3 * so["iterations"])
    assert rel(u2 - u1, dx) <= 1e-4
    # the increment is smaller than the total: the committed criterion is
    # the stricter one
    assert s2["xinf"] < np.abs(u2).max()


def test_simulate_sample_deck(tmp_path):
    run = tmp_path / "beam"
    shutil.copytree(SAMPLE, run)
    u, st, frames = E.simulate(run / "beam.json")
    print("simulate", {k: st[k] for k in ("iterations", "rinf", "xinf", "status", "converged")}, flush=True)
    assert st["status"] == 0 and len(frames) == 2
    assert frames[0].name.endswith("-undeformed.vtk") and frames[1].name.endswith("-deformed.vtk")
    assert all(p.parent.parent == run / "results" for p in frames)
    m, s = E.Mesh.from_input(run / "beam.json")
    g0, g1 = read_legacy_vtk(frames[0]), read_legacy_vtk(frames[1])
    assert np.array_equal(g0["points"].astype(np.float32), m.X.astype(np.float32))
    assert np.array_equal(g1["points"].astype(np.float32), (m.X + scatter(m, u)).astype(np.float32))
    assert g1["cells"] == m.conn.tolist()
    # Integrator::integral's return: ||dx||_inf below dx_criterion (1e-8)
    assert st["converged"] == (st["xinf"] < 1e-8)
    # the driver's solve is the plain solve of the deck (deterministic)
    with E.Context(m) as c:
        ud, sd = c.solve(**s)
    assert np.array_equal(u, ud) and sd["iterations"] == st["iterations"]
    # save_frame false: nothing written
    shutil.rmtree(run / "results")
    _, _, none = E.simulate(run / "beam.json", save_frame=False)
    assert none == [] and not (run / "results").exists()
