"""GPU: the solve loop's look-ahead (the next direction and first trial run
on the device while the host reads the current trial's scalars, and are
discarded when the trial is rejected or the solve stops) must give the same
iterates as the plain loop, bit for bit: same displacement, statistics and
per-iteration trace, across every way an iteration can end -- accepted,
rejected trials, convergence, the iteration cap, the distance target, the
FP64 switch of the mixed mode -- and every two-loop / preconditioner path."""
import numpy as np
import pytest

import efem_amd as E

pytestmark = pytest.mark.gpu

KEYS = ("iterations", "r_evals", "ls_trials", "rinf", "xinf", "target_reached", "fp64_switch_iteration", "status")


def both(mesh, precision=E.FP64, options=(), **kw):
    out = []
    with E.Context(mesh, precision) as c:
        for name, v in options:
            c.set_option(name, v)
        for la in (0, 1):
            c.set_option("lookahead", la)
            u, st = c.solve(**kw)
            out.append((u, st, c.trace()))
    return out


def check_same(res, what):
    (u0, s0, t0), (u1, s1, t1) = res
    print(what, {k: s0[k] for k in ("iterations", "r_evals", "ls_trials")}, flush=True)
    assert np.array_equal(u0, u1), what
    for k in KEYS:
        assert s0[k] == s1[k] or (np.isnan(s0[k]) and np.isnan(s1[k])), (what, k, s0[k], s1[k])
    assert np.array_equal(t0, t1), what
    return t0


def test_lookahead_identical_iterates():
    m = E.Mesh.workload(1, [20, 4, 4, 0, 300.0])
    t = check_same(both(m, epsilon=1e-8, max_iterations=3000), "converged FP64")
    assert t.shape[0] > 100
    check_same(both(m, epsilon=1e-14, epsilon_rel=0.0, max_iterations=37), "iteration cap")
    check_same(both(m, options=[("two_loop_compact", 0)], epsilon=1e-8, max_iterations=3000), "reference order")
    check_same(both(m, options=[("h0_graph", 1)], epsilon=1e-8, max_iterations=3000), "captured graph")
    check_same(both(m, epsilon=1e-8, max_iterations=400, use_precond=False), "H0 = I / theta")
    # the distance target stops the solve early
    with E.Context(m) as c:
        ut, _ = c.solve(epsilon=1e-10, max_iterations=3000)
    check_same(both(m, epsilon=1e-14, epsilon_rel=0.0, max_iterations=3000, x_target=ut, target_tol=1e-4),
               "target")


def test_lookahead_rejected_trials():
    # a load beyond what the reference's line search handles well: early
    # iterations reject first trials
    m = E.Mesh.workload(1, [20, 4, 4, 0, 3000.0])
    t = check_same(both(m, epsilon=1e-8, max_iterations=150), "rejections")
    assert t[:, 2].max() > 1, "no rejected trial: the case does not exercise the roll-back"
    # the cap on trials (the last trial is taken)
    check_same(both(m, epsilon=1e-8, max_iterations=60, max_linesearch=1), "one trial")


def test_lookahead_mixed_fp64_switch():
    m = E.Mesh.workload(2, [30, 6, 6])
    t = check_same(both(m, E.MIXED, epsilon=1e-7, max_iterations=3000, fp64_switch=1e-3), "mixed switch")
    assert t.shape[0] > 10


@pytest.mark.parametrize("precision", [E.FP64, E.MIXED])
def test_second_loop_reads_sweep_output_through_permutation(precision):
    # with plain preconditioner launches the second loop reads the backward
    # sweep's solution through the inverse permutation; with the captured
    # graph it reads the unpermuted copy: the same values in the same order,
    # so the iterates are identical
    m = E.Mesh.workload(1, [20, 4, 4, 0, 300.0])
    res = []
    for graph in (0, 1):
        with E.Context(m, precision) as c:
            c.set_option("h0_graph", graph)
            u, st = c.solve(epsilon=1e-8, max_iterations=3000, fp64_switch=1e-3 if precision == E.MIXED else 0.0)
            res.append((u, st, c.trace()))
    t = check_same(res, f"permuted read vs graph copy, precision {precision}")
    assert t.shape[0] > 100
