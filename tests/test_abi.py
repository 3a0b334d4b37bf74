"""The C-ABI library: loads, exports every declared entry point, and its
host-only parts (workload generators, host factorisation + the device's
block-inverse solve algorithm) run without a GPU."""
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import efem_amd as E
import oracle as O

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "efem_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(efem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = E.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


@pytest.mark.parametrize("which,sizes", [
    (1, (525, 320, 8, 0, 1500, 80)),
    (2, (88641, 480000, 4, 0, 264600, 8000)),
    (3, (37152, 7680, 20, 1, 109440, 1920)),
    (4, (106425, 73728, 10, 0, 316800, 768)),
])
def test_workload_sizes(which, sizes):
    m = E.Mesh.workload(which)
    got = (m.n_nodes, m.n_elems, m.nen, m.reduced, m.ndof, m.faces.shape[0])
    assert got == sizes


@pytest.mark.parametrize("which", [1, 2, 3, 4])
def test_workload_elements_positively_oriented(which):
    m = E.Mesh.workload(which, [8, 4, 4] if which != 1 else None)
    X, c = m.X, m.conn
    if m.nen in (4, 10):
        v = np.einsum("ij,ij->i", np.cross(X[c[:, 1]] - X[c[:, 0]], X[c[:, 2]] - X[c[:, 0]]), X[c[:, 3]] - X[c[:, 0]])
    else:
        v = np.einsum("ij,ij->i", np.cross(X[c[:, 1]] - X[c[:, 0]], X[c[:, 3]] - X[c[:, 0]]), X[c[:, 4]] - X[c[:, 0]])
    assert (v > 0).all()


def test_cylinder_midside_nodes_on_arcs():
    m = E.Mesh.workload(3, [12, 2, 4])
    r = np.hypot(m.X[:, 0], m.X[:, 1])
    levels = np.unique(np.round(r, 12))
    assert np.allclose(levels, np.linspace(0.010, 0.015, 5))


def test_c4_per_element_modulus_distribution():
    m = E.Mesh.workload(4)
    z = np.log(m.params[:, 0] / 114000.0) / 0.2
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    # the pressure window x in [0.02, 0.06] on y = min
    xs = m.X[m.faces.ravel(), 0]
    assert xs.min() >= 0.02 - 1e-12 and xs.max() <= 0.06 + 1e-12
    assert np.allclose(m.X[m.faces.ravel(), 1], 0.0)


@pytest.mark.parametrize("opts", [
    [6, 3, 3, 0.006, 0.003, 0.003, 8, 0, 100.0],
    [6, 3, 3, 0.006, 0.003, 0.003, 4, 1, 100.0, 0.1, 3],
    [4, 2, 2, 0.004, 0.002, 0.002, 10, 0, 100.0],
    [4, 2, 2, 0.004, 0.002, 0.002, 20, 2, 100.0, 0.0, 1, 1],
])
def test_host_factor_and_block_inverse_solve(opts, rng):
    m = E.Mesh.workload(0, opts)
    K = O.Model(m).K(scalar=False)
    b = rng.standard_normal(m.ndof)
    x_ref = spla.spsolve(K.tocsc(), b)
    x, info = E.host_factor_solve(m, b)
    assert np.linalg.norm(x - x_ref) <= 1e-11 * np.linalg.norm(x_ref)
    assert info["nnz"] > 0 and info["levels"] >= 1


def test_host_factor_wide_fronts_backward_stable(rng):
    # 35k DOF: the top separators are 17 x 17 nodes, so the dense front
    # kernels run their blocked paths (several 32-column blocks, partial
    # 8 x 6 register tiles) and the task-parallel split (nr^2 w > 2e7).
    # A backward-stable Cholesky solve leaves |Kx - b| ~ eps |K| |x|.
    m = E.Mesh.workload(2, [40, 16, 16])
    K = O.Model(m).K(scalar=False).tocsr()
    b = rng.standard_normal(m.ndof)
    x, info = E.host_factor_solve(m, b)
    r = K @ x - b
    assert np.linalg.norm(r) <= 1e-13 * spla.norm(K, 1) * np.linalg.norm(x)
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(b)


_KERNEL_PROBE = """
import sys, numpy as np, scipy.sparse.linalg as spla
import efem_amd as E, oracle as O
m = E.Mesh.workload(2, [24, 12, 12])
K = O.Model(m).K(scalar=False).tocsr()
b = np.random.default_rng(5).standard_normal(m.ndof)
x, _ = E.host_factor_solve(m, b)
r = np.linalg.norm(K @ x - b) / (spla.norm(K, 1) * np.linalg.norm(x))
print(E.host_dense_kernel(), r)
sys.exit(0 if r < 1e-13 else 1)
"""


@pytest.mark.parametrize("cap", ["avx2", "portable"])
def test_host_factor_narrower_micro_kernels(cap):
    # the AVX2 and portable register tiles (chosen on CPUs without AVX-512,
    # or forced here) give a backward-stable solve too
    import os
    import subprocess
    import sys
    root = Path(__file__).resolve().parents[1]
    env = dict(os.environ, EFEM_DENSE_KERNEL=cap, PYTHONPATH=str(root))
    out = subprocess.run([sys.executable, "-c", _KERNEL_PROBE], env=env, cwd=root, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    name = out.stdout.split()[0]
    if cap == "portable":
        assert name == "portable"
    else:
        assert name in ("avx2", "portable")  # portable only where the CPU lacks AVX2


def test_host_factor_large_beam_fill_is_modest():
    # a 60-cell slice of the 265k-DOF beam: nested dissection keeps the
    # scalar factor within ~1.1x the oracle's simplicial fill (7.86M)
    m = E.Mesh.workload(2, [60, 20, 20])
    b = np.ones(m.ndof)
    x, info = E.host_factor_solve(m, b)
    assert np.isfinite(x).all()
    assert info["nnz"] < 1.2 * 7.86e6
    assert info["levels"] < 40


def test_create_is_clean_without_device():
    # on a host without a GPU efem_create reports EFEM_ERR_HIP; with one it
    # succeeds.  Either way nothing crashes and the error text is set.
    m = E.Mesh.workload(1)
    try:
        ctx = E.Context(m)
    except E.EfemError as e:
        assert "HIP" in str(e) or "device" in str(e)
    else:
        assert ctx.ndof == 1500
        ctx.close()


def test_create_rejects_bad_input():
    m = E.Mesh.workload(1)
    bad = E.Mesh(m.X, m.conn[:, :7], 0, m.params)
    with pytest.raises(E.EfemError):
        E.Context(bad)
