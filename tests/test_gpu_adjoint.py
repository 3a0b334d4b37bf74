"""GPU: true tangent (Hessian-vector products), Krylov and Newton solves,
strain energy, adjoint gradients and the torch.autograd binding.

Pins:
  * the tangent against central differences of the oracle's residual (an
    independent CPU restatement), <= 1e-6 relative (FP64);
  * the transpose product by the identity w . (K v) = v . (K^T w);
  * the element part's symmetry (no pressure faces: K = K^T bitwise);
  * Krylov solutions by their tangent residual;
  * the energy by dW/du . v = f_int(u) . v (oracle internal force);
  * adjoint gradients against central differences of J through Newton
    equilibrium solves (16 seeded directions on the small bar, relative
    1e-5), and the mixed adjoint against the FP64 one.
"""
import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import face_objective, rel, small, smooth_u, with_per_element_params

pytestmark = pytest.mark.gpu

TCASES = [(8, 0, 0), (8, 1, 1), (4, 1, 0), (10, 0, 0), (10, 2, 1), (20, 2, 1)]


def directed_mesh():
    m0 = small(8, 0, 0)
    return E.Mesh(m0.X, m0.conn, m0.law, m0.params, m0.fixed, m0.faces, 0.0, 0,
                  pdir=np.array([0.0, 4500.0, -2000.0]))


def random_dir(m, rng):
    v = smooth_u(m, rng, 0.05)
    return v / np.abs(v).max() * 1e-5 + 1e-6 * rng.standard_normal(v.shape)


@pytest.mark.parametrize("kind,law,reduced", TCASES)
def test_tangent_matches_oracle_differences(kind, law, reduced, rng):
    m = small(kind, law, reduced)
    print(f"tangent kind={kind} law={law} reduced={reduced} ndof={m.ndof}", flush=True)
    om = O.Model(m)
    with E.Context(m, E.FP64) as ctx:
        u = smooth_u(m, rng, 0.05)
        v = random_dir(m, rng)
        h = 1e-2
        fd = (om.residual(u + h * v) - om.residual(u - h * v)) / (2 * h)
        Kv = ctx.tangent(u, v)
        assert rel(Kv, fd) <= 1e-6, rel(Kv, fd)
    with E.Context(m, E.MIXED) as ctx:
        assert rel(ctx.tangent(u, v), Kv) <= 1e-5


def test_tangent_directed_traction(rng):
    m = directed_mesh()
    om = O.Model(m)
    with E.Context(m) as ctx:
        u = smooth_u(m, rng, 0.05)
        v = random_dir(m, rng)
        h = 1e-2
        fd = (om.residual(u + h * v) - om.residual(u - h * v)) / (2 * h)
        assert rel(ctx.tangent(u, v), fd) <= 1e-6


@pytest.mark.parametrize("which", ["follower", "directed", "c3d20r"])
def test_tangent_transpose_identity(which, rng):
    m = {"follower": lambda: small(8, 0, 0), "directed": directed_mesh,
         "c3d20r": lambda: small(20, 2, 1)}[which]()
    with E.Context(m) as ctx:
        u = smooth_u(m, rng, 0.1)
        v = rng.standard_normal(m.ndof)
        w = rng.standard_normal(m.ndof)
        Kv = ctx.tangent(u, v)
        KTw = ctx.tangent(u, w, transpose=True)
        a, b = w @ Kv, v @ KTw
        assert abs(a - b) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(Kv)
        # the follower load stiffness is not symmetric
        assert rel(ctx.tangent(u, v, transpose=True), Kv) > 1e-9


def test_tangent_symmetric_without_pressure(rng):
    m = small(10, 1, 0, p=0.0)
    assert m.faces.shape[0] == 0
    with E.Context(m) as ctx:
        u = smooth_u(m, rng, 0.1)
        v = rng.standard_normal(m.ndof)
        w = rng.standard_normal(m.ndof)
        Kv = ctx.tangent(u, v)
        assert np.array_equal(Kv, ctx.tangent(u, v, transpose=True))
        a, b = w @ Kv, v @ ctx.tangent(u, w)
        assert abs(a - b) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(Kv)


@pytest.mark.parametrize("p,method", [(0.0, E.KRYLOV_CG), (800.0, E.KRYLOV_BICGSTAB), (800.0, E.KRYLOV_AUTO)])
def test_tangent_solve(p, method, rng):
    m = small(10, 0, 0, p=p)
    with E.Context(m) as ctx:
        u = smooth_u(m, rng, 0.02)
        b = rng.standard_normal(m.ndof)
        for tr in (False, True):
            x, st = ctx.tangent_solve(u, b, transpose=tr, tol=1e-10, method=method)
            assert st["converged"], st
            assert rel(ctx.tangent(u, x, transpose=tr), b) <= 1e-8
            x2, st2 = ctx.tangent_solve(u, b, transpose=tr, tol=1e-10, method=method, use_precond=False,
                                        max_iterations=5000)
            if st2["converged"]:
                assert rel(x2, x) <= 1e-6
            # the Cholesky preconditioner pays
            assert st["iterations"] < st2["iterations"]


def test_newton_reaches_lbfgs_solution():
    m = small(8, 0, 0, p=800.0)
    om = O.Model(m)
    with E.Context(m) as ctx:
        fext = np.abs(ctx.residual(np.zeros(m.ndof))).max()
        un, st = ctx.newton(epsilon=1e-11 * fext)
        assert st["status"] == 0 and st["rinf"] <= 1e-11 * fext and st["iterations"] <= 20, st
        ul, sl = ctx.solve(epsilon=1e-8 * fext, epsilon_rel=0.0, max_iterations=50000)
        assert rel(ul, un) <= 1e-5
        assert np.abs(om.residual(un)).max() <= 1e-10 * fext


def test_energy_derivative_is_internal_force(rng):
    m = small(10, 1, 0, p=0.0)
    om = O.Model(m)
    with E.Context(m) as ctx:
        assert ctx.energy(np.zeros(m.ndof)) == 0.0
        u = smooth_u(m, rng, 0.05)
        v = random_dir(m, rng)
        h = 1e-2
        dW = (ctx.energy(u + h * v) - ctx.energy(u - h * v)) / (2 * h)
        want = om.internal_force(u) @ v
        assert abs(dW - want) <= 1e-6 * abs(want), (dW, want)
        assert ctx.energy(u) == ctx.energy(u)
    with E.Context(m, E.MIXED) as ctx:
        # the energy is always FP64
        W64 = E.Context(m).energy(u)
        assert ctx.energy(u) == W64


def c4_small(rng):
    # the C3D10 bar with per-element C10 at 8x2x2 cells and 1 kPa (20 kPa bends
    # a bar this slender by more than its length)
    m = E.Mesh.workload(4, [8, 2, 2, 7, 1000.0])
    assert m.params_per_elem == 1 and m.faces.shape[0] > 0
    return m


def _equilibrium(ctx, m, C10, p0, z, s, u0, eps):
    P = m.params.copy()
    P[:, 0] = C10 * np.exp(z)
    ctx.set_params(P)
    ctx.set_load(pressure=p0 * np.exp(s))
    u, st = ctx.newton(u0, epsilon=eps, max_iterations=40, tol=1e-10)
    assert st["status"] == 0 and st["rinf"] <= eps, st
    return u


def test_adjoint_matches_central_differences(rng):
    m = c4_small(rng)
    print(f"adjoint bar: {m.n_elems} C3D10, ndof {m.ndof}", flush=True)
    J, g = face_objective(m)
    C10 = m.params[:, 0].copy()
    p0 = m.pressure
    with E.Context(m) as ctx:
        fext = np.abs(ctx.residual(np.zeros(m.ndof))).max()
        eps = 1e-12 * fext
        u0 = _equilibrium(ctx, m, C10, p0, 0.0, 0.0, None, eps)
        lam, gp, gl, st = ctx.adjoint(u0, g(u0), tol=1e-12)
        assert st["converged"], st
        gz = gp[:, 0] * C10          # d J / d log C10_e
        gs = gl[0] * p0              # d J / d log p
        h = 1e-6
        errs = []
        for k in range(16):
            dz = rng.standard_normal(m.n_elems)
            ds = rng.standard_normal()
            up = _equilibrium(ctx, m, C10, p0, h * dz, h * ds, u0, eps)
            um = _equilibrium(ctx, m, C10, p0, -h * dz, -h * ds, u0, eps)
            fd = (J(up) - J(um)) / (2 * h)
            an = gz @ dz + gs * ds
            errs.append(abs(fd - an) / abs(an))
        print("adjoint vs FD relative errors:", np.array(errs), flush=True)
        assert max(errs) <= 1e-5, errs


def test_mixed_adjoint_matches_fp64(rng):
    m = c4_small(rng)
    J, g = face_objective(m)
    with E.Context(m) as c64:
        fext = np.abs(c64.residual(np.zeros(m.ndof))).max()
        u, st = c64.newton(epsilon=1e-12 * fext)
        _, gp64, gl64, _ = c64.adjoint(u, g(u), tol=1e-12)
    with E.Context(m, E.MIXED) as c32:
        _, gp32, gl32, st = c32.adjoint(u, g(u), tol=1e-10)
        assert st["converged"], st
    err_p = rel(gp32[:, 0], gp64[:, 0])
    err_l = abs(gl32[0] - gl64[0]) / abs(gl64[0])
    print(f"mixed vs FP64 adjoint: params {err_p:.2e}, load {err_l:.2e}, {st}", flush=True)
    # FP32-tangent Krylov inside FP64 iterative refinement: FP64 accuracy
    assert err_p <= 1e-7 and err_l <= 1e-7


def test_adjoint_shared_parameters_and_directed_load(rng):
    # shared parameters: the gradient is the sum over elements; directed
    # traction: d J / d pdir; both against central differences
    m = directed_mesh()
    J, g = face_objective(m, target=(0.0, 0.0, -0.0005))
    with E.Context(m) as ctx:
        fext = np.abs(ctx.residual(np.zeros(m.ndof))).max()
        eps = 1e-12 * fext
        u0, _ = ctx.newton(epsilon=eps)
        _, gp, gl, st = ctx.adjoint(u0, g(u0), tol=1e-12)
        assert st["converged"]
        h = 1e-6
        base = np.array(m.params, dtype=np.float64)
        for k in (0, 3):  # C10, D1
            def Jk(sg):
                P = base.copy()
                P[k] *= np.exp(sg * h)
                ctx.set_params(P)
                u, st = ctx.newton(u0, epsilon=eps)
                return J(u)
            fd = (Jk(1.0) - Jk(-1.0)) / (2 * h)
            an = gp[k] * base[k]
            assert abs(fd - an) <= 1e-5 * abs(an) + 1e-14, (k, fd, an)
        ctx.set_params(base)
        for c in range(3):
            def Jc(sg):
                d = m.pdir.copy()
                d[c] += sg * h * 1000.0
                ctx.set_load(pdir=d)
                u, st = ctx.newton(u0, epsilon=eps)
                return J(u)
            fd = (Jc(1.0) - Jc(-1.0)) / (2 * h * 1000.0)
            assert abs(fd - gl[c]) <= 1e-5 * np.abs(gl).max(), (c, fd, gl[c])


def test_torch_autograd_gradcheck(rng):
    torch = pytest.importorskip("torch")
    from efem_amd.autograd import solve_equilibrium
    m0 = E.Mesh.workload(0, [3, 1, 1, 0.003, 0.001, 0.001, 8, 0, 800.0, 0.0, 5, 0])
    m = with_per_element_params(m0, rng)
    J, g = face_objective(m, target=(0.0, 0.0, -1e-5))
    fid = E.free_ids(m.fixed)
    with E.Context(m) as ctx:
        fext = np.abs(ctx.residual(np.zeros(m.ndof))).max()
        C10 = torch.tensor(m.params[:, 0])
        rest = torch.tensor(m.params[:, 1:])
        p0 = m.pressure
        _, gJ = J, g

        def f(z, s):
            params = torch.cat([(C10 * torch.exp(z))[:, None], rest], dim=1)
            u = solve_equilibrium(ctx, params, p0 * torch.exp(s), epsilon=1e-12 * fext)
            return 1e10 * (u ** 2).sum()

        z = torch.zeros(m.n_elems, dtype=torch.float64, requires_grad=True)
        s = torch.zeros((), dtype=torch.float64, requires_grad=True)
        assert torch.autograd.gradcheck(f, (z, s), eps=1e-6, atol=1e-5, rtol=1e-4)
