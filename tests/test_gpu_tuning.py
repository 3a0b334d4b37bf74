"""GPU: the autotuner and the launch-shape options change summation orders
only: residuals agree to rounding, solves reach the same solution, each
shape is deterministic, and invalid shapes are refused."""
import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import rel, small, smooth_u

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,opts", [(2, [40, 8, 8]), (4, [16, 4, 4])])
def test_autotune_keeps_results(which, opts, rng):
    m = E.Mesh.workload(which, opts)
    om = O.Model(m)
    u = smooth_u(m, rng, 0.05)
    with E.Context(m) as c:
        r0 = c.residual(u)
        rep = c.autotune()
        print("autotune", rep, flush=True)
        assert rep["gather_lanes"] in (1, 2, 4, 8) and rep["reduce_blocks"] in (32, 64, 128, 256)
        assert rep["history_blocks"] in (32, 64, 128, 256) and rep["history_ms"] > 0
        r1 = c.residual(u)
        assert rel(r1, om.residual(u)) <= 1e-10 and rel(r1, r0) <= 1e-12
        assert np.array_equal(r1, c.residual(u))


def test_solve_after_autotune():
    # the tuner's timing passes use the history buffers and scalar slots as
    # scratch: a solve afterwards starts from an empty history and reaches
    # the untuned context's solution (a load the reference converges on)
    m = E.Mesh.workload(1, [20, 4, 4, 0, 300.0])
    with E.Context(m) as c:
        ref, st0 = c.solve(epsilon=1e-9, max_iterations=5000)
    assert st0["rinf"] <= 1e-9
    with E.Context(m) as c:
        rep = c.autotune()
        print("autotune", rep, flush=True)
        u, st = c.solve(epsilon=1e-9, max_iterations=5000)
    assert st["rinf"] <= 1e-9 and rel(u, ref) <= 1e-6


@pytest.mark.parametrize("lanes,tpb,blocks,hist", [(1, 64, 32, 256), (2, 128, 64, 32), (8, 256, 256, 128)])
def test_launch_shapes(lanes, tpb, blocks, hist, rng):
    m = E.Mesh.workload(1, [20, 4, 4, 0, 300.0])
    u = smooth_u(m, rng, 0.05)
    with E.Context(m) as c:
        ref, _ = c.solve(epsilon=1e-9, max_iterations=5000)
        r0 = c.residual(u)
        c.set_option("gather_lanes", lanes)
        c.set_option("tet4_threads", tpb)
        c.set_option("reduce_blocks", blocks)
        c.set_option("history_blocks", hist)
        assert rel(c.residual(u), r0) <= 1e-12
        uu, st = c.solve(epsilon=1e-9, max_iterations=5000)
        assert st["rinf"] <= 1e-9 and rel(uu, ref) <= 1e-6
    m4 = E.Mesh.workload(2, [20, 4, 4])
    with E.Context(m4) as c:
        r0 = c.residual(smooth_u(m4, rng, 0.05))
    u4 = smooth_u(m4, rng, 0.05)
    with E.Context(m4) as c:
        a = c.residual(u4)
        c.set_option("tet4_threads", tpb)
        assert np.array_equal(c.residual(u4), a)  # block size: same per-element arithmetic


def test_invalid_shapes_refused():
    m = small(8, 0, 0)
    with E.Context(m) as c:
        for name, v in (("gather_lanes", 3), ("tet4_threads", 96), ("reduce_blocks", 0), ("reduce_blocks", 1000),
                        ("history_blocks", 0), ("history_blocks", 1000)):
            with pytest.raises(E.EfemError):
                c.set_option(name, v)


@pytest.mark.parametrize("prec,tol", [(E.FP64, 1e-10), (E.MIXED, 1e-4)])
@pytest.mark.parametrize("which,opts", [(2, [40, 8, 8]), (3, [24, 2, 8]), (1, None)])
def test_autotune_factor_keeps_the_preconditioner(prec, tol, which, opts, rng):
    # every dense-subtree size is the same K^-1 up to rounding: the tuned
    # context is factored, its apply solves K, and a refactorisation keeps
    # the tuned structure
    import scipy.sparse.linalg as spla
    m = E.Mesh.workload(which, opts)
    K = O.Model(m).K(scalar=False)
    b = rng.standard_normal(m.ndof)
    x_ref = spla.spsolve(K.tocsc(), b)
    with E.Context(m, prec) as c:
        rep = c.autotune_factor()
        print("autotune_factor", rep, flush=True)
        assert rep["dense_subtree"] in (64, 128, 192, 256) and rep["apply_ms"] > 0
        assert c.factor_nnz == rep["factor_nnz"] > 0
        assert rel(c.apply_h0(b), x_ref) <= tol
        assert c.factor()
        assert c.factor_nnz == rep["factor_nnz"]
        assert rel(c.apply_h0(b), x_ref) <= tol


def test_autotune_factor_refused_on_batches():
    ms = [E.Mesh.workload(5, [i, 3, 1.0 / 64]) for i in range(2)]
    with E.Batch(ms, E.FP64) as b:
        with pytest.raises(E.EfemError):
            b.autotune_factor()
