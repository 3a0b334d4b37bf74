"""Shared helpers of the GPU tests (meshes, smooth displacement fields)."""
import numpy as np

import efem_amd as E


def small(kind, law, reduced, p=800.0, jitter=0.05):
    """5x3x2 box of `kind` elements, x=0 fixed, follower pressure p on z=max
    (p = 0: no pressure faces)."""
    m = E.Mesh.workload(0, [5, 3, 2, 0.005, 0.003, 0.002, kind, law, p if p > 0 else 1.0,
                            jitter if kind in (4, 10) else 0.0, 11, reduced])
    if p > 0:
        return m
    return E.Mesh(m.X, m.conn, m.law, m.params, m.fixed, None, 0.0, m.reduced)


def smooth_u(m, rng, amp=0.02):
    """Smooth bending + stretching field plus small noise, zero on fixed nodes."""
    X = m.X
    L = X.max(0) - X.min(0)
    s = X[:, 0] / L[0]
    u = np.stack([amp * L[0] * (0.3 * s + 0.1 * np.sin(3 * s)),
                  amp * L[0] * 0.2 * s ** 2,
                  -amp * L[0] * s ** 2], axis=1)
    u += amp * 0.02 * L.min() * rng.standard_normal(u.shape)
    return u[m.fixed == 0].ravel()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def with_per_element_params(m, rng, sigma=0.2):
    """Copy of mesh m with per-element C10 = C10 exp(sigma z_e)."""
    p = np.tile(np.asarray(m.params, dtype=np.float64).reshape(-1, 4)[:1], (m.n_elems, 1))
    p[:, 0] *= np.exp(sigma * rng.standard_normal(m.n_elems))
    return E.Mesh(m.X, m.conn, m.law, p, m.fixed, m.faces, m.pressure, m.reduced, pdir=m.pdir)


def face_objective(m, target=(0.0, 0.005, 0.0)):
    """J(u) = sum over the free nodes of the x = max face of |u_i - target|^2
    and its gradient dJ/du."""
    xmax = m.X[:, 0].max()
    S = np.flatnonzero(np.isclose(m.X[:, 0], xmax) & (m.fixed == 0))
    fid = E.free_ids(m.fixed)
    idx = (3 * fid[S][:, None] + np.arange(3)).ravel()
    t = np.tile(np.asarray(target, dtype=np.float64), len(S))

    def J(u):
        return float(((u[idx] - t) ** 2).sum())

    def g(u):
        out = np.zeros(len(u))
        out[idx] = 2.0 * (u[idx] - t)
        return out

    return J, g


class DeviceBuffer:
    """A device allocation made through the HIP runtime the library itself
    links (libamdhip64 of /opt/rocm), for tests of the _d entry points.
    torch ships a HIP runtime of its own, which must not share the process's
    device with it, so torch tensors are not used for this."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            E.lib()  # loads the library and with it its HIP runtime
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, n):
        import ctypes as C
        self.n = int(n)
        self.ptr = C.c_void_p()
        rc = self.hip().hipMalloc(C.byref(self.ptr), max(8 * self.n, 8))
        assert rc == 0, rc

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) == 0  # host -> device

    def download(self):
        out = np.empty(self.n)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device -> host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
