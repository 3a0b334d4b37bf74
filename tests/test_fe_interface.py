"""The reference's C++ FE interfaces (include/efem/fe.hpp) over the C ABI:
Material / NeoHookean / MooneyRivlin / Yeoh (src/material/Material.h:9-51),
the Constrainable family and the Constraint aggregate (src/fe/Constraint.h:
23-77) and Integrator::integral (src/simulation/Integrator.cpp:22-42),
exercised by the host program tests/cpp/fe_driver (built by build())."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import efem_amd as E
import oracle as O

DRIVER = Path(__file__).resolve().parent / "cpp" / "fe_driver"


def run(*args, timeout=120):
    if not DRIVER.exists():
        pytest.fail("tests/cpp/fe_driver is not built (run __graft_entry__.build())")
    out = subprocess.run([str(DRIVER), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stdout + out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_materials_match_oracle():
    # getFirstPiola of each class built by initFromJson against the oracle's
    # P(F) (pinned by the reference's own material sources), and the
    # classes' elasticity_coeff against the coefficient the device factor
    # uses (efem_material_point's law, same parameters)
    d = run("materials")
    F = np.array(d["F"]).reshape(3, 3)
    cases = {"P_nh": (E.NEOHOOKEAN, [1.5e5, 0, 0, 2e-6]),
             "P_mr": (E.MOONEY_RIVLIN, [1.2e5, 3e4, 0, 5e-6]),
             "P_yeoh": (E.YEOH, [1e5, -2e3, 150, 1e-5])}
    for key, (law, p) in cases.items():
        P = np.array(d[key]).reshape(3, 3)
        want = O.piola(law, p, F)
        assert np.abs(P - want).max() <= 1e-12 * np.abs(want).max(), key
    # the oracle's coefficient for the same parameters: the oracle's K of a
    # one-tet model scales with it, so K(law, p) = coeff x K(NH with
    # C10 = 0, D1 = 2), whose coefficient is 2 / 2 = 1
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    conn = np.array([[0, 1, 2, 3]], np.int32)
    for i, (key, (law, p)) in enumerate(cases.items()):
        m = E.Mesh(X, conn, law, np.array(p, float), np.zeros(4, np.uint8), np.zeros((0, 3), np.int32), 0.0, 0)
        K1 = O.Model(m).K(scalar=True).toarray()
        m0 = E.Mesh(X, conn, 0, np.array([0.0, 0, 0, 2.0]), np.zeros(4, np.uint8), np.zeros((0, 3), np.int32), 0.0, 0)
        K0 = O.Model(m0).K(scalar=True).toarray()
        assert np.allclose(K1, d["coeff"][i] * K0, rtol=1e-14, atol=0), key
    assert d["ids"] == [3, 4, 5] and d["density"] == 1100
    assert d["types"] == [3, 4, 5]  # IN_NEOHOOKEAN, IN_MOONEY_RIVLIN, IN_YEOH


def test_unrepresentable_problems_are_refused():
    d = run("refusals")
    assert d == {"no_device_law": 1, "mixed_laws": 1, "no_elements": 1, "two_pressures": 1,
                 "two_force_vectors": 1}


@pytest.mark.parametrize("which,opts", [(1, [10, 2, 2, 0, 300.0]), (4, [6, 2, 2])])
def test_host_jacobi_matches_oracle_K(which, opts):
    # efem_host_jacobi (ElementConstraint::conGetJacobi's matrix) against the
    # oracle's K, and its 3 x 3 block pattern with explicit zeros
    import scipy.sparse as sp
    m = E.Mesh.workload(which, opts)
    r, c, v = E.host_jacobi(m)
    n = m.ndof
    K = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsc()
    Ko = O.Model(m).K(scalar=False).tocsc()
    assert abs(K - Ko).max() <= 1e-12 * abs(Ko).max()
    assert len(v) == 9 * O.Model(m).K(scalar=True).nnz
    off = (r % 3) != (c % 3)
    assert np.all(v[off] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("which,opts", [(1, [10, 2, 2, 0, 300.0]), (2, [12, 3, 3, 1, 100.0])])
def test_constraint_aggregate_on_device(which, opts):
    # the classes' residual equals efem_residual on the same problem (one
    # device path), the constraints' parts sum to it, the Jacobian triplets
    # are efem_host_jacobi's, integral() is efem_solve, and after updateDx
    # the committed state is in equilibrium
    print(f"fe_driver solve {which} {opts}", flush=True)
    d = run("solve", which, *opts, timeout=300)
    assert "error" not in d, d
    assert d["res_vs_abi"] == 0.0
    assert d["parts_vs_total"] <= 1e-12 * d["res_scale"]
    assert d["jacobi_mismatch"] == 0 and d["jacobi_nnz"] > 0
    assert d["dx_vs_abi"] == 0.0 and d["iterations"] == d["abi_iterations"] > 0
    assert d["converged"] == 1
    # the committed state is the solve's last iterate: its residual is the
    # solve's final one, and a second integral() starts from it
    assert d["committed_vs_dx"] == 0.0
    assert d["committed_residual"] == d["rinf"]
    assert d["second_rinf"] <= d["rinf"]
