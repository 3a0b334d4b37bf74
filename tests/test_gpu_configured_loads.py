"""GPU: the reference algorithm at the configured loads, device FP64 against
the oracle's record (tests/golden/configured_load_traces.json, written by
tests/golden/make_configured_load_traces.py with the single-thread oracle:
LBFGS.h:96-174, projection line search LineSearchFEA.h:37-68, m = 6, one
load step from u = 0 as Simulator.cpp:20-41 runs it).

At the cantilever's 2 kPa the reference iteration does not converge, and at
the beam's 3 kPa and the bar's 20 kPa it diverges.  The device's FP64 solve
takes the oracle's first ten iterations (same trial counts and steps, the
same |r|inf to 1e-6) and ends in the same state.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import efem_amd as E

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "configured_load_traces.json").read_text())["cases"]


@pytest.mark.parametrize("name", sorted(GOLD))
def test_configured_load_first_iterations_and_status(name):
    g = GOLD[name]
    m = E.Mesh.workload(g["workload"], g["opts"])
    print(f"configured load: {name}, ndof {m.ndof}", flush=True)
    ref = np.array(g["trace"])
    with E.Context(m, E.FP64) as ctx:
        _, st = ctx.solve(epsilon=1e-6, max_iterations=10)
        tr = ctx.trace()
        assert st["iterations"] == 10
        assert np.array_equal(tr[:, 1:], ref[:10, 1:])
        assert np.allclose(tr[:, 0], ref[:10, 0], rtol=1e-6, atol=0)
        # the whole capped run: the oracle's status
        u, st = ctx.solve(epsilon=1e-6, max_iterations=g["iterations"])
        trf = ctx.trace()
    rmax = np.nanmax(trf[:, 0]) if st["status"] == 0 else np.inf
    converged = st["status"] == 0 and (st["rinf"] <= 1e-6 or st["rinf"] <= 1e-6 * st["xinf"])
    diverged = st["status"] != 0 or not np.isfinite(rmax) or rmax > 1e3 * g["r0_inf"]
    status = "converged" if converged else ("diverged" if diverged else "no convergence")
    print(f"  device {status} after {st['iterations']} it; oracle {g['status']} after {g['iterations']}", flush=True)
    assert status == g["status"]
