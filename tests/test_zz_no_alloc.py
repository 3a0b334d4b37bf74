"""GPU: the allocation-free energy entry point (efem_energy_noalloc) against
efem_energy, through the C ABI: same kernels and summation order, so the
same value bit for bit, on two calls in a row (the second on a smaller
problem, so the scratch array holds the first call's larger contents), and
refusal of a problem too large for the scratch."""
import ctypes as C

import numpy as np
import pytest

import efem_amd as E
from helpers import DeviceBuffer, smooth_u

pytestmark = pytest.mark.gpu


def _energy_noalloc(ctx, u):
    ud, wd = DeviceBuffer(ctx.ndof), DeviceBuffer(1)
    try:
        ud.upload(u)
        wd.upload(np.array([np.nan]))
        rc = ctx._lib.efem_energy_noalloc(ctx._h, C.c_void_p(ud.ptr.value), C.c_void_p(wd.ptr.value))
        return rc, wd.download()[0]
    finally:
        ud.free()
        wd.free()


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_energy_noalloc_matches_energy_on_two_calls(prec, rng):
    big = E.Mesh.workload(4, [32, 8, 6])    # C3D10 bar, per-element C10
    small = E.Mesh.workload(1, [10, 2, 2, 0, 300.0])  # C3D8 cantilever
    for m in (big, small):
        print(f"noalloc energy nen={m.nen} elems={m.n_elems} ndof={m.ndof} prec={prec}", flush=True)
        u = smooth_u(m, rng, 0.05)
        with E.Context(m, prec) as ctx:
            rc, w = _energy_noalloc(ctx, u)
            assert rc == E.EFEM_OK, ctx._lib.efem_last_error(ctx._h)
            want = ctx.energy(u)
            assert np.isfinite(want) and want > 0
            assert w == want, (w, want)
            # and again on the same context: the scratch is reused
            rc, w2 = _energy_noalloc(ctx, 0.5 * u)
            assert rc == E.EFEM_OK and w2 == ctx.energy(0.5 * u)


def test_energy_noalloc_refuses_problems_beyond_the_scratch():
    # 2^21 doubles of scratch: the full beam (480,000 elements, 264,600 DOF)
    # fits, a 1.6 M-element beam does not; the refusal touches no device memory
    m = E.Mesh.workload(2, [200, 20, 20])
    with E.Context(m, E.FP64) as ctx:
        rc, w = _energy_noalloc(ctx, np.zeros(ctx.ndof))
        assert rc == E.EFEM_OK and w == 0.0
    m = E.Mesh.workload(2, [300, 30, 30])
    with E.Context(m, E.FP64) as ctx:
        ud, wd = DeviceBuffer(1), DeviceBuffer(1)
        try:
            rc = ctx._lib.efem_energy_noalloc(ctx._h, C.c_void_p(ud.ptr.value), C.c_void_p(wd.ptr.value))
        finally:
            ud.free()
            wd.free()
        assert rc == E.EFEM_ERR_ARG
        assert b"too large" in ctx._lib.efem_last_error(ctx._h)
