"""Device tests of the reference's failure semantics, through the C ABI.

* a non-finite result zeroes the displacement and reports failure
  (src/simulation/Integrator.cpp:28-34): EFEM_ERR_NONFINITE, u = 0;
* a factorisation that fails (K not SPD, or no free DOF) falls back to
  H0 = I/theta (src/simulation/nlsolver/LBFGS.h:71-80,
  LBFGSMat.h:119-123): precond_failed = 1 and the iterates follow the
  oracle's fallback path;
* invalid solver parameters are refused (check_param,
  src/simulation/nlsolver/Param.h:110-127): EFEM_ERR_ARG for each;
* a batch context refuses per-context load / parameter changes
  (EFEM_ERR_STATE), and efem_residual_d on device buffers equals
  efem_residual.
"""
import ctypes as C

import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import DeviceBuffer, rel, smooth_u

pytestmark = pytest.mark.gpu


def _params(**kw):
    base = dict(epsilon=1e-6, epsilon_rel=1e-6, max_iterations=10, max_linesearch=20, m=6, lns_fea=1.0,
                use_precond=1, line_search=0, x_target=None, target_tol=0.0, fp64_switch=0.0)
    base.update(kw)
    return E.LbfgsParams(base["epsilon"], base["epsilon_rel"], base["max_iterations"], base["max_linesearch"],
                         base["m"], base["lns_fea"], base["use_precond"], base["line_search"], base["x_target"],
                         base["target_tol"], base["fp64_switch"])


def _raw_solve(ctx, u, prm):
    st = E.SolveStats()
    rc = ctx._lib.efem_solve(ctx._h, u.ctypes.data_as(C.POINTER(C.c_double)), C.byref(prm), C.byref(st))
    return rc, st


def test_nonfinite_displacement_is_zeroed_and_reported():
    m = E.Mesh.workload(1, [10, 2, 2, 0, 100.0])
    print("case nan-in-u cantilever 10x2x2", flush=True)
    with E.Context(m, E.FP64) as ctx:
        u0 = np.zeros(ctx.ndof)
        u0[7] = np.nan
        u, st = ctx.solve(u0=u0, epsilon=1e-6, max_iterations=50)
        assert st["status"] == E.EFEM_ERR_NONFINITE
        assert np.array_equal(u, np.zeros(ctx.ndof))
        assert not np.isfinite(st["rinf"])
        # the context stays usable: a clean solve afterwards converges
        u2, st2 = ctx.solve(epsilon=1e-6, max_iterations=2000)
        assert st2["status"] == 0 and st2["rinf"] <= 1e-6 and np.all(np.isfinite(u2))


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_indefinite_jacobian_falls_back_to_theta(prec):
    # elasticity_coeff = 2.94001 C10 + 2/D1 < 0 (C10 = -2000, D1 = 1e-2):
    # K = c L is negative definite, the Cholesky factorisation fails, and the
    # solve runs with H0 = I/theta (first direction -r), as the oracle does
    m0 = E.Mesh.workload(1, [10, 2, 2, 0, 1.0])
    p = np.array([-2000.0, 0.0, 0.0, 1e-2])
    m = E.Mesh(m0.X, m0.conn, m0.law, p, m0.fixed, m0.faces, m0.pressure, m0.reduced)
    om = O.Model(m)
    assert not om.factor(scalar=True)
    u_or, st_or = om.solve(epsilon=0.0, max_iterations=6, trace_cap=6)
    assert st_or["precond_failed"] == 1
    with E.Context(m, prec) as ctx:
        assert not ctx.factor()
        u, st = ctx.solve(epsilon=0.0, max_iterations=6)
        tr = ctx.trace()
    print("fallback gpu", st, "oracle", st_or, flush=True)
    assert st["precond_failed"] == 1 and st["factor_nnz"] == 0
    assert st["iterations"] == st_or["iterations"] == 6
    assert st["status"] == 0 and np.all(np.isfinite(u))
    if prec == E.FP64:
        # the first direction is -r: its line search expands the step past
        # element inversion (finite stress, as in the reference) and back
        assert np.array_equal(tr[:, 2], st_or["trace"][:, 2])
        assert np.allclose(tr[:, 0], st_or["trace"][:, 0], rtol=1e-8)
        assert rel(u, u_or) <= 1e-8


def test_all_fixed_mesh_has_no_factor():
    # every node fixed: no free DOF, K is empty; the reference's
    # jacobi.nonZeros() == 0 branch (LBFGS.h:71-80) -> fallback
    m0 = E.Mesh.workload(1, [4, 1, 1, 0, 100.0])
    m = E.Mesh(m0.X, m0.conn, m0.law, m0.params, np.ones_like(m0.fixed), None, 0.0, m0.reduced)
    with E.Context(m, E.FP64) as ctx:
        assert ctx.ndof == 0
        assert not ctx.factor()
        u, st = ctx.solve(epsilon=1e-6, max_iterations=5)
        assert st["status"] == 0 and st["iterations"] == 1 and u.size == 0


@pytest.mark.parametrize("field,value", [
    ("epsilon", -1e-6), ("epsilon_rel", -1.0), ("max_iterations", -1), ("max_linesearch", 0),
    ("max_linesearch", -3), ("m", 0), ("m", -2), ("m", 65), ("line_search", 7),
])
def test_invalid_solver_parameters_are_refused(field, value):
    m = E.Mesh.workload(1, [6, 2, 2, 0, 100.0])
    with E.Context(m, E.FP64) as ctx:
        u = np.zeros(ctx.ndof)
        rc, _ = _raw_solve(ctx, u, _params(**{field: value}))
        assert rc == E.EFEM_ERR_ARG, (field, value, rc)
        assert b"invalid" in ctx._lib.efem_last_error(ctx._h) or b"line-search" in ctx._lib.efem_last_error(ctx._h)
        # the refusal leaves u untouched and the context usable
        assert np.array_equal(u, np.zeros(ctx.ndof))
        rc, st = _raw_solve(ctx, u, _params())
        assert rc == E.EFEM_OK and st.iterations == 10


def test_batch_refuses_per_context_load_and_parameter_changes():
    meshes = [E.Mesh.workload(5, [i, 7, 1.0 / 64]) for i in range(3)]
    with E.Batch(meshes, E.FP64) as b:
        rc = b._lib.efem_set_load(b._h, 500.0, None)
        assert rc == E.EFEM_ERR_STATE
        p = np.ascontiguousarray(np.atleast_2d(meshes[0].params)[0], dtype=np.float64)
        rc = b._lib.efem_set_params(b._h, p.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == E.EFEM_ERR_STATE
        # the batch still solves as created
        u, st = b.solve(epsilon=1e-6, max_iterations=3000)
        oms = [O.Model(mm) for mm in meshes]
        off = b.offsets
        for i, om in enumerate(oms):
            ui = u[off[i]:off[i + 1]]
            assert np.abs(om.residual(ui)).max() <= 1e-5


def test_residual_on_device_buffers_matches_host_entry(rng):
    m = E.Mesh.workload(2, [40, 8, 8])
    u = smooth_u(m, rng, 0.05)
    with E.Context(m, E.FP64) as ctx:
        r_host = ctx.residual(u)
        ud, rd = DeviceBuffer(ctx.ndof), DeviceBuffer(ctx.ndof)
        try:
            ud.upload(u)
            rd.upload(np.full(ctx.ndof, np.nan))
            ctx.residual_device(ud.ptr.value, rd.ptr.value)
            # efem_residual_d returns with r_d written: read it at once
            assert np.array_equal(rd.download(), r_host)
            # u_d may be overwritten right after the call
            ud.upload(2.0 * u)
            ctx.residual_device(ud.ptr.value, rd.ptr.value)
            assert np.array_equal(rd.download(), ctx.residual(2.0 * u))
        finally:
            ud.free()
            rd.free()
