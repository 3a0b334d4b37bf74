"""efem_amd -- MI355X (gfx950) Total-Lagrangian solve loop.

Python binding of the C ABI declared in include/efem_gpu.h (libefem_gpu.so,
built in-tree by efem_amd/_build.py).  The library is required: importing
this package without it raises, and every compute call goes to the HIP
kernels -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libefem_gpu.so"

EFEM_OK = 0
EFEM_ERR_ARG = -1
EFEM_ERR_HIP = -2
EFEM_ERR_NONFINITE = -3
EFEM_ERR_STATE = -4

NEOHOOKEAN, MOONEY_RIVLIN, YEOH = 0, 1, 2
FP64, F32_ELEMENTS, F32_PRECOND, MIXED = 0, 1, 2, 3
LS_PROJECTION, LS_ENERGY = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class Problem(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int), ("X", _dp), ("n_elems", C.c_int), ("nen", C.c_int),
        ("conn", _ip), ("reduced", C.c_int), ("law", C.c_int), ("params", _dp),
        ("params_per_elem", C.c_int), ("fixed", _up), ("n_faces", C.c_int), ("fnn", C.c_int),
        ("faces", _ip), ("pressure", C.c_double), ("pdir", _dp), ("n_fnodes", C.c_int),
        ("fnodes", _ip), ("fvec", _dp), ("volumes", _dp), ("precision", C.c_int),
        ("gravity", C.c_int), ("consistent_inertia", C.c_int), ("density", _dp),
    ]


class LbfgsParams(C.Structure):
    _fields_ = [
        ("epsilon", C.c_double), ("epsilon_rel", C.c_double), ("max_iterations", C.c_int),
        ("max_linesearch", C.c_int), ("m", C.c_int), ("lns_fea", C.c_double),
        ("use_precond", C.c_int), ("line_search", C.c_int), ("x_target", _dp),
        ("target_tol", C.c_double), ("fp64_switch", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("r_evals", C.c_int), ("ls_trials", C.c_int),
        ("precond_failed", C.c_int), ("rinf", C.c_double), ("xinf", C.c_double),
        ("factor_seconds", C.c_double), ("solve_seconds", C.c_double),
        ("factor_nnz", C.c_longlong), ("target_reached", C.c_int), ("target_err", C.c_double),
        ("fp64_switch_iteration", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


KRYLOV_AUTO, KRYLOV_CG, KRYLOV_BICGSTAB = 0, 1, 2


class KrylovParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iterations", C.c_int), ("method", C.c_int),
                ("use_precond", C.c_int)]


class KrylovStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("precond_applies", C.c_int),
                ("converged", C.c_int), ("rel_residual", C.c_double), ("seconds", C.c_double),
                ("factor_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def load_library(path: Path | str | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is not built; run `python efem_amd/_build.py` (or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    sig = {
        "efem_create": (C.c_int, [C.POINTER(vp), C.POINTER(Problem)]),
        "efem_destroy": (C.c_int, [vp]),
        "efem_last_error": (C.c_char_p, [vp]),
        "efem_global_error": (C.c_char_p, []),
        "efem_ndof": (C.c_int, [vp]),
        "efem_precision": (C.c_int, [vp]),
        "efem_residual": (C.c_int, [vp, _dp, _dp]),
        "efem_residual_d": (C.c_int, [vp, vp, vp]),
        "efem_factor": (C.c_int, [vp, _ip]),
        "efem_factor_nnz": (C.c_longlong, [vp]),
        "efem_factor_layout": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "efem_factor_release": (C.c_int, [vp]),
        "efem_factor_phases": (C.c_int, [vp, _dp]),
        "efem_profile_read_h0": (C.c_int, [vp, _dp, _ip]),
        "efem_apply_h0": (C.c_int, [vp, _dp, _dp]),
        "efem_solve": (C.c_int, [vp, _dp, C.POINTER(LbfgsParams), C.POINTER(SolveStats)]),
        "efem_trace": (C.c_int, [vp, _dp, C.c_int]),
        "efem_set_option": (C.c_int, [vp, C.c_char_p, C.c_double]),
        "efem_host_factor_solve": (C.c_int, [C.POINTER(Problem), _dp, _dp, C.POINTER(C.c_longlong)]),
        "efem_host_dense_kernel": (C.c_char_p, []),
        "efem_host_constant_force": (C.c_int, [C.POINTER(Problem), _dp, C.c_int]),
        "efem_host_jacobi": (C.c_int, [C.POINTER(Problem), C.POINTER(C.c_longlong), _ip, _ip, _dp, C.c_longlong]),
        "efem_device_memory": (C.c_int, [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "efem_commit": (C.c_int, [vp, _dp]),
        "efem_committed": (C.c_int, [vp, _dp]),
        "efem_write_vtk": (C.c_int, [C.POINTER(Problem), _dp, C.c_char_p]),
        "efem_results_path": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
        "efem_profile_reset": (C.c_int, [vp]),
        "efem_profile_read": (C.c_int, [vp, _dp, _ip]),
        "efem_tangent": (C.c_int, [vp, _dp, _dp, _dp, C.c_int]),
        "efem_tangent_solve": (C.c_int, [vp, _dp, _dp, _dp, C.c_int, C.POINTER(KrylovParams),
                                         C.POINTER(KrylovStats)]),
        "efem_adjoint": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, C.POINTER(KrylovParams),
                                   C.POINTER(KrylovStats)]),
        "efem_newton": (C.c_int, [vp, _dp, C.c_double, C.c_int, C.POINTER(KrylovParams),
                                  C.POINTER(SolveStats)]),
        "efem_energy": (C.c_int, [vp, _dp, _dp]),
        "efem_energy_noalloc": (C.c_int, [vp, vp, vp]),
        "efem_time_residual": (C.c_int, [vp, _dp, C.c_int, _dp]),
        "efem_autotune": (C.c_int, [vp, _dp]),
        "efem_autotune_factor": (C.c_int, [vp, _dp]),
        "efem_set_params": (C.c_int, [vp, _dp]),
        "efem_set_load": (C.c_int, [vp, C.c_double, _dp]),
        "efem_material_point": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "efem_batch_create": (C.c_int, [C.POINTER(vp), C.POINTER(Problem), C.c_int, C.c_int]),
        "efem_batch_count": (C.c_int, [vp]),
        "efem_batch_offsets": (C.c_int, [vp, _ip]),
        "efem_batch_solve": (C.c_int, [vp, _dp, C.POINTER(LbfgsParams), C.POINTER(SolveStats)]),
        "efem_input_load": (vp, [C.c_char_p]),
        "efem_input_error": (C.c_char_p, []),
        "efem_input_free": (None, [vp]),
        "efem_input_sizes": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "efem_input_problem": (C.c_int, [vp, C.POINTER(Problem), C.c_int]),
        "efem_input_lbfgs_params": (C.c_int, [vp, C.POINTER(LbfgsParams)]),
        "efem_mesh_build": (vp, [C.c_int, _dp, C.c_int]),
        "efem_mesh_free": (None, [vp]),
        "efem_mesh_sizes": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "efem_mesh_arrays": (C.c_int, [vp, _dp, _ip, _dp, _up, _ip, _dp]),
        "efem_mesh_problem": (C.c_int, [vp, C.POINTER(Problem), C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


class Mesh:
    """A problem description held as numpy arrays (reference node order).

    Fields mirror efem_problem: X [n,3], conn [E,nen], reduced, law,
    params [4] or [E,4], fixed [n] (uint8), faces [F,fnn], pressure,
    optional pdir [3], fnodes / fvec (constant nodal forces), volumes [E];
    gravity (0 none, 1..6 = +x, -x, +y, -y, +z, -z), consistent_inertia and
    density ([1] or [E], indexed as params).
    """

    def __init__(self, X, conn, law, params, fixed=None, faces=None, pressure=0.0, reduced=0,
                 pdir=None, fnodes=None, fvec=None, volumes=None, gravity=0, consistent_inertia=0,
                 density=None):
        self.X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.nen = int(self.conn.shape[1])
        self.law = int(law)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self.params_per_elem = 1 if p.ndim == 2 else 0
        self.params = p
        n = self.X.shape[0]
        self.fixed = (np.zeros(n, np.uint8) if fixed is None
                      else np.ascontiguousarray(fixed, dtype=np.uint8))
        self.faces = (np.zeros((0, 3), np.int32) if faces is None
                      else np.ascontiguousarray(faces, dtype=np.int32))
        self.pressure = float(pressure)
        self.reduced = int(reduced)
        self.pdir = None if pdir is None else np.ascontiguousarray(pdir, dtype=np.float64)
        self.fnodes = None if fnodes is None else np.ascontiguousarray(fnodes, dtype=np.int32)
        self.fvec = None if fvec is None else np.ascontiguousarray(fvec, dtype=np.float64)
        self.volumes = None if volumes is None else np.ascontiguousarray(volumes, dtype=np.float64)
        self.gravity = int(gravity)
        self.consistent_inertia = int(consistent_inertia)
        self.density = None if density is None else np.ascontiguousarray(np.atleast_1d(density), dtype=np.float64)

    @property
    def n_nodes(self):
        return self.X.shape[0]

    @property
    def n_elems(self):
        return self.conn.shape[0]

    @property
    def n_free(self):
        return int((self.fixed == 0).sum())

    @property
    def ndof(self):
        return 3 * self.n_free

    @staticmethod
    def workload(which: int, opts=None) -> "Mesh":
        """Synthetic workloads generated by the C++ host code
        (1 cantilever C3D8, 2 beam Kuhn C3D4, 3 cylinder C3D20R,
        4 bar Kuhn C3D10 with per-element C10, 0 custom box)."""
        L = lib()
        o = None if opts is None else np.ascontiguousarray(opts, dtype=np.float64)
        h = L.efem_mesh_build(which, _ptr(o, C.c_double), 0 if o is None else len(o))
        if not h:
            raise RuntimeError("efem_mesh_build failed")
        try:
            s = (C.c_longlong * 8)()
            L.efem_mesh_sizes(h, s)
            n, E, nen, red, law, ppe, F, fnn = (int(v) for v in s)
            X = np.empty((n, 3))
            conn = np.empty((E, nen), np.int32)
            params = np.empty((E, 4) if ppe else (4,))
            fixed = np.empty(n, np.uint8)
            faces = np.empty((F, fnn), np.int32)
            pr = C.c_double()
            L.efem_mesh_arrays(h, _ptr(X, C.c_double), _ptr(conn, C.c_int), _ptr(params, C.c_double),
                               _ptr(fixed, C.c_ubyte), _ptr(faces, C.c_int), C.byref(pr))
        finally:
            L.efem_mesh_free(h)
        return Mesh(X, conn, law, params, fixed, faces, pr.value, red)

    @staticmethod
    def from_input(json_path) -> tuple["Mesh", dict]:
        """An EFEM run directory (JSON config + mesh / constraint files) in
        the reference's formats.  Returns the mesh and the solver settings
        of its analysis block (keyword arguments of Context.solve)."""
        L = lib()
        h = L.efem_input_load(str(json_path).encode())
        if not h:
            raise EfemError(f"efem_input_load: {L.efem_input_error().decode()}")
        try:
            s = (C.c_longlong * 10)()
            L.efem_input_sizes(h, s)
            n, E, nen, law, ppe, F, fnn, nfix, nfn, skipped = (int(v) for v in s)
            pb = Problem()
            L.efem_input_problem(h, C.byref(pb), FP64)
            X = np.ctypeslib.as_array(pb.X, (n * 3,)).reshape(n, 3).copy()
            conn = np.ctypeslib.as_array(pb.conn, (E * nen,)).reshape(E, nen).copy()
            params = np.ctypeslib.as_array(pb.params, (E * 4 if ppe else 4,)).copy()
            if ppe:
                params = params.reshape(E, 4)
            fixed = np.ctypeslib.as_array(pb.fixed, (n,)).copy()
            faces = (np.ctypeslib.as_array(pb.faces, (F * fnn,)).reshape(F, fnn).copy() if F else None)
            pdir = np.ctypeslib.as_array(pb.pdir, (3,)).copy() if pb.pdir else None
            fnodes = np.ctypeslib.as_array(pb.fnodes, (nfn,)).copy() if nfn else None
            fvec = np.ctypeslib.as_array(pb.fvec, (3,)).copy() if nfn else None
            gravity, consistent = int(pb.gravity), int(pb.consistent_inertia)
            density = (np.ctypeslib.as_array(pb.density, (E if ppe else 1,)).copy() if pb.density else None)
            prm = LbfgsParams()
            L.efem_input_lbfgs_params(h, C.byref(prm))
        finally:
            L.efem_input_free(h)
        mesh = Mesh(X, conn, law, params, fixed, faces, pb.pressure, 0, pdir=pdir, fnodes=fnodes, fvec=fvec,
                    gravity=gravity, consistent_inertia=consistent, density=density)
        mesh.skipped_elements = skipped
        settings = {"epsilon": prm.epsilon, "epsilon_rel": prm.epsilon_rel, "max_iterations": prm.max_iterations,
                    "max_linesearch": prm.max_linesearch, "m": prm.m, "lns_fea": prm.lns_fea,
                    "use_precond": bool(prm.use_precond)}
        return mesh, settings

    def problem(self, precision: int = FP64) -> Problem:
        pb = Problem()
        pb.n_nodes = self.n_nodes
        pb.X = _ptr(self.X, C.c_double)
        pb.n_elems = self.n_elems
        pb.nen = self.nen
        pb.conn = _ptr(self.conn, C.c_int)
        pb.reduced = self.reduced
        pb.law = self.law
        pb.params = _ptr(self.params, C.c_double)
        pb.params_per_elem = self.params_per_elem
        pb.fixed = _ptr(self.fixed, C.c_ubyte)
        pb.n_faces = int(self.faces.shape[0])
        pb.fnn = int(self.faces.shape[1]) if self.faces.ndim == 2 else 0
        pb.faces = _ptr(self.faces, C.c_int) if self.faces.size else None
        pb.pressure = self.pressure
        pb.pdir = _ptr(self.pdir, C.c_double)
        pb.n_fnodes = 0 if self.fnodes is None else len(self.fnodes)
        pb.fnodes = _ptr(self.fnodes, C.c_int)
        pb.fvec = _ptr(self.fvec, C.c_double)
        pb.volumes = _ptr(self.volumes, C.c_double)
        pb.precision = int(precision)
        pb.gravity = self.gravity
        pb.consistent_inertia = self.consistent_inertia
        pb.density = _ptr(self.density, C.c_double)
        return pb


class EfemError(RuntimeError):
    pass


class Context:
    """One device context (efem_create ... efem_destroy)."""

    def __init__(self, mesh: Mesh, precision: int = FP64):
        self._lib = lib()
        self.mesh = mesh
        self._pb = mesh.problem(precision)
        h = C.c_void_p()
        rc = self._lib.efem_create(C.byref(h), C.byref(self._pb))
        if rc != EFEM_OK:
            raise EfemError(f"efem_create failed ({rc}): {self._lib.efem_global_error().decode()}")
        self._h = h
        self.ndof = self._lib.efem_ndof(h)

    def _check(self, rc, what):
        if rc != EFEM_OK:
            raise EfemError(f"{what} failed ({rc}): {self._lib.efem_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.efem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def residual(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.empty(self.ndof)
        self._check(self._lib.efem_residual(self._h, _ptr(u, C.c_double), _ptr(r, C.c_double)),
                    "efem_residual")
        return r

    def residual_device(self, u_ptr: int, r_ptr: int):
        self._check(self._lib.efem_residual_d(self._h, C.c_void_p(u_ptr), C.c_void_p(r_ptr)),
                    "efem_residual_d")

    def factor(self) -> bool:
        failed = C.c_int(0)
        self._check(self._lib.efem_factor(self._h, C.byref(failed)), "efem_factor")
        return failed.value == 0

    @property
    def factor_nnz(self) -> int:
        return int(self._lib.efem_factor_nnz(self._h))

    @property
    def factor_layout(self) -> dict:
        """The factor's apply: stored entries, those the level sweeps read,
        the fused top (columns, Z's leading dimension, depth), launches per
        apply inside the solve loop, levels, supernodes."""
        out = (C.c_longlong * 8)()
        self._check(self._lib.efem_factor_layout(self._h, out), "efem_factor_layout")
        keys = ("nnz", "sweep_entries", "top_columns", "top_ld", "top_depth", "launches", "levels", "supernodes")
        return {k: int(v) for k, v in zip(keys, out)}

    def apply_h0(self, v) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty(self.ndof)
        self._check(self._lib.efem_apply_h0(self._h, _ptr(v, C.c_double), _ptr(out, C.c_double)),
                    "efem_apply_h0")
        return out

    def solve(self, u0=None, epsilon=1e-6, epsilon_rel=None, max_iterations=0, max_linesearch=20,
              m=6, lns_fea=1.0, use_precond=True, x_target=None, target_tol=0.0, fp64_switch=0.0,
              line_search=LS_PROJECTION):
        """LBFGSSolver::solve.  epsilon_rel defaults to epsilon, as
        Integrator.cpp:13-14 sets both from residual_criterion.  x_target /
        target_tol: also stop once ||x - x_target|| <= tol ||x_target||."""
        u = np.zeros(self.ndof) if u0 is None else np.array(u0, dtype=np.float64)
        xt = None if x_target is None else np.ascontiguousarray(x_target, dtype=np.float64)
        prm = LbfgsParams(float(epsilon), float(epsilon if epsilon_rel is None else epsilon_rel),
                          int(max_iterations), int(max_linesearch), int(m), float(lns_fea),
                          1 if use_precond else 0, int(line_search), _ptr(xt, C.c_double), float(target_tol),
                          float(fp64_switch))
        st = SolveStats()
        rc = self._lib.efem_solve(self._h, _ptr(u, C.c_double), C.byref(prm), C.byref(st))
        if rc not in (EFEM_OK, EFEM_ERR_NONFINITE):
            self._check(rc, "efem_solve")
        d = st.as_dict()
        d["status"] = rc
        return u, d

    def trace(self) -> np.ndarray:
        n = self._lib.efem_trace(self._h, None, 0)
        out = np.empty((max(n, 0), 3))
        if n > 0:
            self._lib.efem_trace(self._h, _ptr(out, C.c_double), n)
        return out

    def profile_reset(self):
        self._check(self._lib.efem_profile_reset(self._h), "efem_profile_reset")

    def profile_read(self):
        ms = C.c_double()
        k = C.c_int()
        self._check(self._lib.efem_profile_read(self._h, C.byref(ms), C.byref(k)), "efem_profile_read")
        return ms.value, k.value

    def profile_read_h0(self):
        ms = C.c_double()
        k = C.c_int()
        self._check(self._lib.efem_profile_read_h0(self._h, C.byref(ms), C.byref(k)), "efem_profile_read_h0")
        return ms.value, k.value

    def time_residual(self, u, count=50) -> float:
        """Mean device milliseconds per residual evaluation (count back to back)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        ms = C.c_double()
        self._check(self._lib.efem_time_residual(self._h, _ptr(u, C.c_double), int(count), C.byref(ms)),
                    "efem_time_residual")
        return ms.value

    def autotune(self) -> dict:
        """Pick the fastest launch shapes for this problem (efem_autotune)."""
        rep = np.zeros(9)
        self._check(self._lib.efem_autotune(self._h, _ptr(rep, C.c_double)), "efem_autotune")
        return {"gather_lanes": int(rep[0]), "tet4_threads": int(rep[1]), "reduce_blocks": int(rep[2]),
                "history_blocks": int(rep[5]), "residual_ms": float(rep[3]), "dot_ms": float(rep[4]),
                "history_ms": float(rep[6]), "batch_threads": int(rep[7]), "batch_ms": float(rep[8])}

    def autotune_factor(self) -> dict:
        """Pick the dense-subtree size with the fastest preconditioner apply
        (efem_autotune_factor); the context stays factored with it."""
        rep = np.zeros(4)
        self._check(self._lib.efem_autotune_factor(self._h, _ptr(rep, C.c_double)), "efem_autotune_factor")
        return {"dense_subtree": int(rep[0]), "apply_ms": float(rep[1]), "factor_s": float(rep[2]),
                "factor_nnz": int(rep[3])}

    def set_option(self, name: str, value: float):
        self._check(self._lib.efem_set_option(self._h, name.encode(), float(value)), "efem_set_option")

    def factor_release(self):
        self._check(self._lib.efem_factor_release(self._h), "efem_factor_release")

    def factor_phases(self) -> dict:
        """Seconds of the last factorisation by phase (efem_factor_phases)."""
        out = np.zeros(8)
        self._check(self._lib.efem_factor_phases(self._h, _ptr(out, C.c_double)), "efem_factor_phases")
        keys = ("host_model", "symbolic", "numeric", "descriptors", "repack", "upload", "other_buffers", "total")
        return dict(zip(keys, out.tolist()))

    def commit(self, u):
        """Constraint::updateDx: u becomes the committed displacement the
        next solve's increment (and its relative criterion) is measured
        from; None clears it."""
        if u is None:
            self._check(self._lib.efem_commit(self._h, None), "efem_commit")
            return
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.ndof,):
            raise ValueError(f"u must have {self.ndof} entries")
        self._check(self._lib.efem_commit(self._h, _ptr(u, C.c_double)), "efem_commit")

    def committed(self) -> np.ndarray:
        out = np.empty(self.ndof)
        self._check(self._lib.efem_committed(self._h, _ptr(out, C.c_double)), "efem_committed")
        return out

    # ---- differentiation -------------------------------------------------
    def tangent(self, u, v, transpose=False) -> np.ndarray:
        """K_t(u) v (or K_t^T v): the true tangent d r / d u applied to v."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty(self.ndof)
        self._check(self._lib.efem_tangent(self._h, _ptr(u, C.c_double), _ptr(v, C.c_double),
                                           _ptr(out, C.c_double), 1 if transpose else 0), "efem_tangent")
        return out

    @staticmethod
    def _kp(tol, max_iterations, method, use_precond):
        return KrylovParams(float(tol), int(max_iterations), int(method), 1 if use_precond else 0)

    def tangent_solve(self, u, b, transpose=False, tol=1e-10, max_iterations=1000, method=KRYLOV_AUTO,
                      use_precond=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.ndof)
        prm = self._kp(tol, max_iterations, method, use_precond)
        st = KrylovStats()
        self._check(self._lib.efem_tangent_solve(self._h, _ptr(u, C.c_double), _ptr(b, C.c_double),
                                                 _ptr(x, C.c_double), 1 if transpose else 0, C.byref(prm),
                                                 C.byref(st)), "efem_tangent_solve")
        return x, st.as_dict()

    def newton(self, u0=None, epsilon=1e-9, max_iterations=50, tol=1e-8, max_krylov=500,
               method=KRYLOV_AUTO, use_precond=True):
        """Newton-Krylov on the true tangent until ||r||_inf <= epsilon."""
        u = np.zeros(self.ndof) if u0 is None else np.array(u0, dtype=np.float64)
        prm = self._kp(tol, max_krylov, method, use_precond)
        st = SolveStats()
        rc = self._lib.efem_newton(self._h, _ptr(u, C.c_double), float(epsilon), int(max_iterations),
                                   C.byref(prm), C.byref(st))
        if rc not in (EFEM_OK, EFEM_ERR_NONFINITE):
            self._check(rc, "efem_newton")
        d = st.as_dict()
        d["status"] = rc
        return u, d

    def adjoint(self, u, g, tol=1e-10, max_iterations=1000, method=KRYLOV_AUTO, use_precond=True):
        """Adjoint gradient at an equilibrium u for dJ/du = g.  Returns
        (lambda, grad_params [E,4] or [4], grad_load [3], stats)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        lam = np.empty(self.ndof)
        gp = np.empty((self.mesh.n_elems, 4) if self.mesh.params_per_elem else (4,))
        gl = np.zeros(3)
        prm = self._kp(tol, max_iterations, method, use_precond)
        st = KrylovStats()
        self._check(self._lib.efem_adjoint(self._h, _ptr(u, C.c_double), _ptr(g, C.c_double),
                                           _ptr(lam, C.c_double), _ptr(gp, C.c_double), _ptr(gl, C.c_double),
                                           C.byref(prm), C.byref(st)), "efem_adjoint")
        return lam, gp, gl, st.as_dict()

    def energy(self, u) -> float:
        u = np.ascontiguousarray(u, dtype=np.float64)
        W = C.c_double()
        self._check(self._lib.efem_energy(self._h, _ptr(u, C.c_double), C.byref(W)), "efem_energy")
        return W.value

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        if p.size != self.mesh.params.size:
            raise ValueError(f"params: expected {self.mesh.params.size} values, got {p.size}")
        self._check(self._lib.efem_set_params(self._h, _ptr(p, C.c_double)), "efem_set_params")

    def set_load(self, pressure=0.0, pdir=None):
        d = None if pdir is None else np.ascontiguousarray(pdir, dtype=np.float64)
        self._check(self._lib.efem_set_load(self._h, float(pressure), _ptr(d, C.c_double)), "efem_set_load")


class Batch(Context):
    """Independent instances in one context (efem_batch_create); solve()
    runs every instance's L-BFGS in one persistent launch."""

    def __init__(self, meshes, precision: int = FP64):
        self._lib = lib()
        self.meshes = list(meshes)
        self.mesh = self.meshes[0]
        pbs = (Problem * len(self.meshes))(*[m.problem(precision) for m in self.meshes])
        self._pbs = pbs
        h = C.c_void_p()
        rc = self._lib.efem_batch_create(C.byref(h), pbs, len(self.meshes), int(precision))
        if rc != EFEM_OK:
            raise EfemError(f"efem_batch_create failed ({rc}): {self._lib.efem_global_error().decode()}")
        self._h = h
        self.ndof = self._lib.efem_ndof(h)
        self.count = self._lib.efem_batch_count(h)
        off = np.empty(self.count + 1, np.int32)
        self._check(self._lib.efem_batch_offsets(h, _ptr(off, C.c_int)), "efem_batch_offsets")
        self.offsets = off.astype(np.int64)

    def split(self, u):
        return [u[self.offsets[i]:self.offsets[i + 1]] for i in range(self.count)]

    def solve(self, u0=None, epsilon=1e-6, epsilon_rel=None, max_iterations=10000, max_linesearch=20, m=6,
              lns_fea=1.0, use_precond=True, fp64_switch=0.0, x_target=None, target_tol=0.0):
        """efem_batch_solve; x_target (concatenated like u) stops each
        instance at its own ||x_i - x_target_i|| <= target_tol ||x_target_i||."""
        u = np.zeros(self.ndof) if u0 is None else np.array(u0, dtype=np.float64)
        xt = None if x_target is None else np.ascontiguousarray(x_target, dtype=np.float64)
        prm = LbfgsParams(float(epsilon), float(epsilon if epsilon_rel is None else epsilon_rel),
                          int(max_iterations), int(max_linesearch), int(m), float(lns_fea),
                          1 if use_precond else 0, 0, _ptr(xt, C.c_double), float(target_tol), float(fp64_switch))
        st = (SolveStats * self.count)()
        rc = self._lib.efem_batch_solve(self._h, _ptr(u, C.c_double), C.byref(prm), st)
        if rc not in (EFEM_OK, EFEM_ERR_NONFINITE):
            self._check(rc, "efem_batch_solve")
        out = [s.as_dict() for s in st]
        for d in out:
            d["status"] = rc
        return u, out


def material_point(law, params, F, dF=None):
    """Host-side P(F), W(F) and, with dF, the tangent product dP = dP/dF [dF]."""
    L = lib()
    p = np.ascontiguousarray(params, dtype=np.float64)
    Fm = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    P = np.empty(9)
    W = C.c_double()
    dFm = None if dF is None else np.ascontiguousarray(dF, dtype=np.float64).reshape(9)
    dP = None if dF is None else np.empty(9)
    rc = L.efem_material_point(int(law), _ptr(p, C.c_double), _ptr(Fm, C.c_double), _ptr(P, C.c_double),
                               C.byref(W), _ptr(dFm, C.c_double), _ptr(dP, C.c_double))
    if rc != EFEM_OK:
        raise EfemError(f"efem_material_point failed ({rc})")
    return P.reshape(3, 3), W.value, (None if dP is None else dP.reshape(3, 3))


def host_factor_solve(mesh: Mesh, b):
    """Host-only: x = K^-1 b through the supernodal factor and the device
    block-inverse algorithm run on the CPU (no GPU needed)."""
    L = lib()
    pb = mesh.problem(FP64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty_like(b)
    info = (C.c_longlong * 8)()
    rc = L.efem_host_factor_solve(C.byref(pb), _ptr(b, C.c_double), _ptr(x, C.c_double), info)
    if rc != EFEM_OK:
        raise EfemError(f"efem_host_factor_solve failed ({rc}): {L.efem_global_error().decode()}")
    keys = ("nnz", "supernodes", "levels", "flops", "microseconds", "numeric_us", "laplacian_us",
            "ordering_us")
    return x, dict(zip(keys, (int(v) for v in info)))


def host_constant_force(mesh: Mesh) -> np.ndarray:
    """Host-only: the constant external force on the free DOFs (nodal forces
    plus gravity M g), as the residual subtracts it."""
    pb = mesh.problem(FP64)
    out = np.zeros(3 * mesh.n_free)
    rc = lib().efem_host_constant_force(C.byref(pb), _ptr(out, C.c_double), len(out))
    if rc != EFEM_OK:
        raise EfemError(f"efem_host_constant_force failed ({rc}): {lib().efem_global_error().decode()}")
    return out


def host_jacobi(mesh: Mesh):
    """Host-only: the approximate Jacobian's triplets (efem_host_jacobi) as
    (rows, cols, vals): every 3 x 3 block of the elasticity-weighted scalar
    Laplacian, explicit zeros off the block diagonals."""
    L = lib()
    pb = mesh.problem(FP64)
    nnz = C.c_longlong()
    rc = L.efem_host_jacobi(C.byref(pb), C.byref(nnz), None, None, None, 0)
    if rc != EFEM_OK:
        raise EfemError(f"efem_host_jacobi failed ({rc}): {L.efem_global_error().decode()}")
    r = np.empty(nnz.value, np.int32)
    c = np.empty(nnz.value, np.int32)
    v = np.empty(nnz.value)
    rc = L.efem_host_jacobi(C.byref(pb), C.byref(nnz), _ptr(r, C.c_int), _ptr(c, C.c_int), _ptr(v, C.c_double),
                            nnz.value)
    if rc != EFEM_OK:
        raise EfemError(f"efem_host_jacobi failed ({rc}): {L.efem_global_error().decode()}")
    return r, c, v


def write_vtk(mesh: Mesh, u, path) -> None:
    """Model::saveVTK's grid at X + u (u over the free DOFs, None for the
    undeformed frame) as an ASCII legacy VTK file."""
    pb = mesh.problem(FP64)
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    if uu is not None and uu.shape != (3 * mesh.n_free,):
        raise ValueError(f"u must have {3 * mesh.n_free} entries")
    rc = lib().efem_write_vtk(C.byref(pb), _ptr(uu, C.c_double), str(path).encode())
    if rc != EFEM_OK:
        raise EfemError(f"efem_write_vtk failed ({rc}): {lib().efem_global_error().decode()}")


def results_path(working_dir, name: str) -> Path:
    """<working_dir>/results/<date>/<date-time-ms>-<name>.vtk (Model.cpp:93-110),
    directories created."""
    buf = C.create_string_buffer(4096)
    rc = lib().efem_results_path(str(working_dir).encode(), name.encode(), buf, len(buf))
    if rc != EFEM_OK:
        raise EfemError(f"efem_results_path failed ({rc}): {lib().efem_global_error().decode()}")
    return Path(buf.value.decode())


def simulate(json_path, precision: int = FP64, save_frame=None):
    """Simulator::initial + Simulator::start (src/simulation/Simulator.cpp:6-40)
    for one run directory: the undeformed frame, one static solve from
    dx = 0 (Integrator::integral), the commit of the result
    (Constraint::updateDx) and the deformed frame.  save_frame defaults to
    the deck's output.save_frame (true when absent, Config.h:43-44).
    Returns the displacement, the solve statistics with "converged"
    (||dx||_inf < dx_criterion, Integrator.cpp:36) and the frames written."""
    import json as _json

    json_path = Path(json_path)
    cfg = _json.loads(json_path.read_text())
    out_cfg = cfg.get("output", {}) if isinstance(cfg.get("output", {}), dict) else {}
    if save_frame is None:
        save_frame = out_cfg.get("save_frame", True) if isinstance(out_cfg.get("save_frame"), bool) else True
    an = cfg.get("analysis", {}) if isinstance(cfg.get("analysis", {}), dict) else {}
    dx_criterion = an.get("dx_criterion", 1e-8)
    if isinstance(dx_criterion, bool) or not isinstance(dx_criterion, (int, float)):
        dx_criterion = 1e-8
    mesh, settings = Mesh.from_input(json_path)
    # IO::IO: the working directory is the config path up to its last '/'
    wd = str(json_path)[: str(json_path).rfind("/") + 1]
    frames = []
    if save_frame:
        p = results_path(wd, "undeformed")
        write_vtk(mesh, None, p)
        frames.append(p)
    with Context(mesh, precision) as ctx:
        u, st = ctx.solve(None, **settings)
        if st["status"] != EFEM_OK:
            # Integrator.cpp:28-34: a non-finite result is dropped
            u = np.zeros_like(u)
        ctx.commit(u)
    st["converged"] = bool(st["status"] == EFEM_OK and st["xinf"] < dx_criterion)
    if save_frame:
        p = results_path(wd, "deformed")
        write_vtk(mesh, u, p)
        frames.append(p)
    return u, st, frames


def device_memory():
    """(free, total) bytes of the current HIP device."""
    f, t = C.c_longlong(0), C.c_longlong(0)
    rc = lib().efem_device_memory(C.byref(f), C.byref(t))
    if rc != EFEM_OK:
        raise EfemError(f"efem_device_memory failed ({rc}): {lib().efem_global_error().decode()}")
    return f.value, t.value


def host_dense_kernel() -> str:
    """Name of the host factorisation's dense micro-kernel."""
    return lib().efem_host_dense_kernel().decode()


def free_ids(fixed) -> np.ndarray:
    """node -> free id (or -1), Constraint.cpp:61-72."""
    fixed = np.asarray(fixed)
    fid = -np.ones(len(fixed), np.int64)
    free = np.flatnonzero(fixed == 0)
    fid[free] = np.arange(len(free))
    return fid
