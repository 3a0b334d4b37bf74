"""GPU: the reference's sample run (sample/beam, copied to tests/golden/
sample_beam) read with efem_input_* and solved with the settings of its own
config (lbfgs_ah, residual_criterion 1e-6, m = 6) -- the same solution as the
oracle's solve of the same deck, and the first iterations identical."""
from pathlib import Path

import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import rel

pytestmark = pytest.mark.gpu

SAMPLE = Path(__file__).resolve().parent / "golden" / "sample_beam"


def test_sample_beam_solves_like_oracle():
    m, s = E.Mesh.from_input(SAMPLE / "beam.json")
    print(f"sample beam: {m.n_elems} C3D4, {m.ndof} DOF, traction {m.pdir}", flush=True)
    om = O.Model(m)
    with E.Context(m) as c:
        u0 = np.zeros(m.ndof)
        assert rel(c.residual(u0), om.residual(u0)) <= 1e-12
        kw = dict(s, max_iterations=20000)
        u, st = c.solve(**kw)
        tr = c.trace()
    print("GPU", {k: st[k] for k in ("iterations", "r_evals", "rinf", "xinf", "status")}, flush=True)
    assert st["status"] == 0 and (st["rinf"] <= s["epsilon"] or st["rinf"] <= s["epsilon"] * st["xinf"])
    uo, so = om.solve(epsilon=s["epsilon"], epsilon_rel=s["epsilon_rel"], max_iterations=20000, trace_cap=10)
    print("oracle", {k: so[k] for k in ("iterations", "rinf")}, flush=True)
    assert np.allclose(tr[:10, 0], so["trace"][:10, 0], rtol=1e-6)
    # both stop at the deck's residual criterion; their displacements agree
    # to the accuracy that criterion allows
    assert rel(u, uo) <= 1e-3
    # tighter: the GPU Newton polish to the rounding floor against the
    # oracle's L-BFGS continued to 1e-9 N
    with E.Context(m) as c:
        ut, sn = c.newton(u, epsilon=1e-14, max_iterations=20)
    uot, sot = om.solve(uo, epsilon=1e-9, epsilon_rel=0.0, max_iterations=50000)
    print("newton", sn["rinf"], "oracle", sot["iterations"], sot["rinf"], flush=True)
    assert sot["rinf"] <= 1e-9
    assert rel(ut, uot) <= 1e-5
