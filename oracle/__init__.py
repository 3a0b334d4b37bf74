"""ctypes binding of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (efem_amd).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_materials.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class OrProblem(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int), ("X", _dp), ("n_elems", C.c_int), ("nen", C.c_int), ("conn", _ip),
        ("reduced", C.c_int), ("law", C.c_int), ("params", _dp), ("params_per_elem", C.c_int),
        ("fixed", _up), ("n_faces", C.c_int), ("fnn", C.c_int), ("faces", _ip),
        ("pressure", C.c_double), ("pdir", _dp), ("n_fnodes", C.c_int), ("fnodes", _ip),
        ("fvec", _dp), ("volumes", _dp), ("precision", C.c_int), ("gravity", C.c_int),
        ("consistent_inertia", C.c_int), ("density", _dp),
    ]


class OrParams(C.Structure):
    _fields_ = [
        ("epsilon", C.c_double), ("epsilon_rel", C.c_double), ("max_iterations", C.c_int),
        ("max_linesearch", C.c_int), ("m", C.c_int), ("lns_fea", C.c_double),
        ("use_precond", C.c_int), ("factor_scalar", C.c_int), ("line_search", C.c_int),
        ("x_target", _dp), ("target_tol", C.c_double),
    ]


class OrStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("r_evals", C.c_int), ("ls_trials", C.c_int),
        ("precond_failed", C.c_int), ("rinf", C.c_double), ("xinf", C.c_double),
        ("factor_seconds", C.c_double), ("solve_seconds", C.c_double),
        ("target_reached", C.c_int), ("target_err", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


RESID_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)
H0_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)
ENERGY_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, _dp, _dp, _dp)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        sig = {
            "or_shape": (C.c_int, [C.c_int, _dp, _dp, _dp]),
            "or_n_gauss": (C.c_int, [C.c_int, C.c_int]),
            "or_gauss": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "or_element_volume": (C.c_double, [C.c_int, _dp]),
            "or_piola": (None, [C.c_int, _dp, _dp, _dp]),
            "or_elasticity_coeff": (C.c_double, [C.c_int, _dp]),
            "or_energy_density": (C.c_double, [C.c_int, _dp, _dp]),
            "or_potential": (C.c_double, [vp, _dp]),
            "or_potential_diff": (C.c_double, [vp, _dp, _dp, _dp]),
            "or_lbfgs_energy": (C.c_int, [C.c_int, RESID_FN, ENERGY_FN, H0_FN, vp, _dp, C.POINTER(OrParams),
                                          C.POINTER(OrStats), _dp, C.c_int]),
            "or_model_create": (vp, [C.POINTER(OrProblem), C.c_char_p, C.c_int]),
            "or_model_destroy": (None, [vp]),
            "or_ndof": (C.c_int, [vp]),
            "or_free_id": (C.c_int, [vp, C.c_int]),
            "or_volume_of": (C.c_double, [vp, C.c_int]),
            "or_residual": (None, [vp, _dp, _dp]),
            "or_internal_force": (None, [vp, _dp, _dp]),
            "or_external_force": (None, [vp, _dp, _dp]),
            "or_face_parent": (C.c_int, [vp, C.c_int]),
            "or_K_nnz": (C.c_long, [vp, C.c_int]),
            "or_K_csc": (None, [vp, C.c_int, _ip, _ip, _dp]),
            "or_factor": (C.c_int, [vp, C.c_int]),
            "or_factor_nnz": (C.c_long, [vp]),
            "or_apply_h0": (None, [vp, _dp, _dp]),
            "or_lbfgs": (C.c_int, [C.c_int, RESID_FN, H0_FN, vp, _dp, C.POINTER(OrParams),
                                   C.POINTER(OrStats), _dp, C.c_int]),
            "or_lbfgs_solve_model": (C.c_int, [vp, _dp, C.POINTER(OrParams), C.POINTER(OrStats),
                                               _dp, C.c_int]),
            "or_two_loop": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double,
                                   H0_FN, vp, _dp, C.c_double, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own material sources (oracle/_ref), or None."""
    global _ref
    if _ref is None and REF_LIB.exists():
        R = C.CDLL(str(REF_LIB))
        R.ref_piola.restype = None
        R.ref_piola.argtypes = [C.c_int, _dp, _dp, _dp]
        R.ref_elasticity_coeff.restype = C.c_double
        R.ref_elasticity_coeff.argtypes = [C.c_int, _dp]
        _ref = R
    return _ref


def _p(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def shape(nen, xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    N = np.zeros(nen)
    dN = np.zeros((nen, 3))
    lib().or_shape(nen, _p(xi), _p(N), _p(dN))
    return N, dN


def gauss(nen, reduced):
    n = lib().or_n_gauss(nen, reduced)
    pts = np.zeros((n, 3))
    w = np.zeros(n)
    for q in range(n):
        xi = np.zeros(3)
        wq = C.c_double()
        lib().or_gauss(nen, reduced, q, _p(xi), C.byref(wq))
        pts[q] = xi
        w[q] = wq.value
    return pts, w


def piola(law, p, F):
    p = np.ascontiguousarray(p, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    P = np.zeros(9)
    lib().or_piola(law, _p(p), _p(F), _p(P))
    return P.reshape(3, 3)


def energy_density(law, p, F):
    p = np.ascontiguousarray(p, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    return lib().or_energy_density(law, _p(p), _p(F))


def ref_piola(law, p, F):
    R = ref_lib()
    p = np.ascontiguousarray(p, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    P = np.zeros(9)
    R.ref_piola(law, _p(p), _p(F), _p(P))
    return P.reshape(3, 3)


def element_volume(nen, Xe):
    Xe = np.ascontiguousarray(Xe, dtype=np.float64)
    return lib().or_element_volume(nen, _p(Xe))


class Model:
    """Oracle model built from an efem_amd.Mesh-like object."""

    def __init__(self, mesh):
        self.mesh = mesh
        pb = OrProblem()
        pb.n_nodes = mesh.X.shape[0]
        pb.X = _p(mesh.X)
        pb.n_elems = mesh.conn.shape[0]
        pb.nen = mesh.conn.shape[1]
        pb.conn = _p(mesh.conn, C.c_int)
        pb.reduced = mesh.reduced
        pb.law = mesh.law
        pb.params = _p(mesh.params)
        pb.params_per_elem = 1 if mesh.params.ndim == 2 else 0
        pb.fixed = _p(mesh.fixed, C.c_ubyte)
        pb.n_faces = mesh.faces.shape[0]
        pb.fnn = mesh.faces.shape[1] if mesh.faces.ndim == 2 else 0
        pb.faces = _p(mesh.faces, C.c_int) if mesh.faces.size else None
        pb.pressure = mesh.pressure
        pb.pdir = _p(mesh.pdir)
        pb.n_fnodes = 0 if mesh.fnodes is None else len(mesh.fnodes)
        pb.fnodes = _p(mesh.fnodes, C.c_int)
        pb.fvec = _p(mesh.fvec)
        pb.volumes = _p(mesh.volumes)
        # gravity (InertiaConstraint): absent on meshes that do not carry it
        pb.gravity = int(getattr(mesh, "gravity", 0))
        pb.consistent_inertia = int(getattr(mesh, "consistent_inertia", 0))
        density = getattr(mesh, "density", None)
        self._density = None if density is None else np.ascontiguousarray(np.atleast_1d(density), dtype=np.float64)
        pb.density = _p(self._density)
        self._pb = pb
        err = C.create_string_buffer(256)
        self._h = lib().or_model_create(C.byref(pb), err, 256)
        if not self._h:
            raise RuntimeError("or_model_create: " + err.value.decode())
        self.ndof = lib().or_ndof(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_model_destroy(self._h)
            self._h = None

    def residual(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.ndof)
        lib().or_residual(self._h, _p(u), _p(r))
        return r

    def internal_force(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.ndof)
        lib().or_internal_force(self._h, _p(u), _p(r))
        return r

    def external_force(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.ndof)
        lib().or_external_force(self._h, _p(u), _p(r))
        return r

    def potential_diff(self, x, xp):
        """(Pi(x) - Pi(xp) element by element, rounding scale at xp)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        xp = np.ascontiguousarray(xp, dtype=np.float64)
        sc = C.c_double()
        d = lib().or_potential_diff(self._h, _p(x), _p(xp), C.byref(sc))
        return d, sc.value

    def potential(self, u):
        """Strain energy minus the constant loads' work (NaN with pressure faces)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().or_potential(self._h, _p(u))

    def volume(self, e):
        return lib().or_volume_of(self._h, e)

    def face_parent(self, f):
        return lib().or_face_parent(self._h, f)

    def K(self, scalar=False):
        import scipy.sparse as sp
        nnz = lib().or_K_nnz(self._h, int(scalar))
        n = self.ndof // 3 if scalar else self.ndof
        Ap = np.zeros(n + 1, np.int32)
        Ai = np.zeros(nnz, np.int32)
        Ax = np.zeros(nnz)
        lib().or_K_csc(self._h, int(scalar), _p(Ap, C.c_int), _p(Ai, C.c_int), _p(Ax))
        return sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))

    def factor(self, scalar=True):
        return lib().or_factor(self._h, int(scalar)) == 0

    def factor_nnz(self):
        return lib().or_factor_nnz(self._h)

    def apply_h0(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.ndof)
        lib().or_apply_h0(self._h, _p(v), _p(out))
        return out

    def solve(self, u0=None, epsilon=1e-6, epsilon_rel=None, max_iterations=0, max_linesearch=20,
              m=6, lns_fea=1.0, use_precond=True, factor_scalar=False, trace_cap=0, line_search=0,
              x_target=None, target_tol=0.0):
        u = np.zeros(self.ndof) if u0 is None else np.array(u0, dtype=np.float64)
        xt = None if x_target is None else np.ascontiguousarray(x_target, dtype=np.float64)
        prm = OrParams(epsilon, epsilon if epsilon_rel is None else epsilon_rel, max_iterations,
                       max_linesearch, m, lns_fea, int(use_precond), int(factor_scalar), int(line_search),
                       _p(xt), float(target_tol))
        st = OrStats()
        tr = np.zeros((max(trace_cap, 1), 3))
        rc = lib().or_lbfgs_solve_model(self._h, _p(u), C.byref(prm), C.byref(st), _p(tr), trace_cap)
        if rc < 0:
            raise ValueError("or_lbfgs_solve_model refused the parameters")
        d = st.as_dict()
        if trace_cap:
            d["trace"] = tr[: min(st.iterations, trace_cap)]
        return u, d


def lbfgs(n, resid, x0, h0=None, epsilon=1e-6, epsilon_rel=None, max_iterations=0,
          max_linesearch=20, m=6, lns_fea=1.0, trace_cap=0, energy=None):
    """Generic oracle L-BFGS on a Python residual callback r = resid(x)."""
    def _rf(ctx, xp, rp):
        x = np.ctypeslib.as_array(xp, shape=(n,))
        r = np.ctypeslib.as_array(rp, shape=(n,))
        r[:] = resid(x.copy())

    def _hf(ctx, vp, op):
        v = np.ctypeslib.as_array(vp, shape=(n,))
        o = np.ctypeslib.as_array(op, shape=(n,))
        o[:] = h0(v.copy())

    def _ef(ctx, xa, xb, scale):
        ea = float(energy(np.ctypeslib.as_array(xa, shape=(n,)).copy()))
        eb = float(energy(np.ctypeslib.as_array(xb, shape=(n,)).copy()))
        scale[0] = abs(eb)
        return ea - eb

    rf = RESID_FN(_rf)
    hf = H0_FN(_hf) if h0 is not None else H0_FN()
    ef = ENERGY_FN(_ef) if energy is not None else ENERGY_FN()
    x = np.array(x0, dtype=np.float64)
    prm = OrParams(epsilon, epsilon if epsilon_rel is None else epsilon_rel, max_iterations,
                   max_linesearch, m, lns_fea, 1 if h0 else 0, 0, 1 if energy is not None else 0, None, 0.0)
    st = OrStats()
    tr = np.zeros((max(trace_cap, 1), 3))
    k = lib().or_lbfgs_energy(n, rf, ef, hf, None, _p(x), C.byref(prm), C.byref(st), _p(tr), trace_cap)
    d = st.as_dict()
    d["return"] = k
    if trace_cap:
        d["trace"] = tr[: min(st.iterations, trace_cap)]
    return x, d


def two_loop(S, Y, ys, theta, ncorr, ptr, v, a=-1.0, h0=None):
    """LBFGSMat::apply_Hv on explicit history (S, Y: [m, n])."""
    mcap, n = S.shape
    S = np.ascontiguousarray(S)
    Y = np.ascontiguousarray(Y)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    res = np.zeros(n)

    def _hf(ctx, vp, op):
        vv = np.ctypeslib.as_array(vp, shape=(n,))
        o = np.ctypeslib.as_array(op, shape=(n,))
        o[:] = h0(vv.copy())

    hf = H0_FN(_hf) if h0 is not None else H0_FN()
    lib().or_two_loop(n, mcap, ncorr, ptr, _p(S), _p(Y), _p(ys), theta, hf, None, _p(v), a, _p(res))
    return res
