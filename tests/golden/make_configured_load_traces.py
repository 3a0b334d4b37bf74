"""Writes tests/golden/configured_load_traces.json: what the reference
algorithm does at the configured loads, from the oracle (the single-thread
C restatement of src/simulation/nlsolver/LBFGS.h:96-174 with the projection
line search of LineSearchFEA.h:37-68, m = 6, max 20 trials, the reference's
stopping test |r|inf <= 1e-6 N or <= 1e-6 |x|inf, approximate-Jacobian H0
factored in the scalar-Laplacian mode).  One load step from u = 0, as
Simulator::start runs (src/simulation/Simulator.cpp:20-41).

For each configured mesh and load: the per-iteration trace
{|r|inf, accepted step, trials} over the first `iterations` iterations, the
initial |r|inf, and a status: "converged" (the stopping test was met),
"diverged" (|r|inf exceeded 1e3 |r0|inf) or "no convergence" (neither
within the cap).  The device's FP64 solve is compared with the first rows
(tests/test_gpu_configured_loads.py); bench.py reports the device's own run
at the configured load beside the oracle's status.

Run from the repository root: python tests/golden/make_configured_load_traces.py
(minutes: the beam's factorisation alone takes ~15 s, each of its iterations
up to 20 residual evaluations of 0.12 s).
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import efem_amd as E  # noqa: E402
import oracle as O  # noqa: E402

CASES = [
    # name, workload, opts, iterations
    ("C1 cantilever 20x4x4 C3D8 NH, 2 kPa", 1, [20, 4, 4, 0, 2000.0], 3000),
    ("C2 beam 200x20x20 Kuhn C3D4 MR, 3 kPa", 2, [200, 20, 20, 20240531, 3000.0], 200),
    ("C4 bar 64x16x12 Kuhn C3D10 NH, 20 kPa", 4, [64, 16, 12, 7, 20000.0], 200),
]


def run(which, opts, iterations):
    m = E.Mesh.workload(which, opts)
    om = O.Model(m)
    r0 = float(np.abs(om.residual(np.zeros(m.ndof))).max())
    t = time.perf_counter()
    u, st = om.solve(epsilon=1e-6, max_iterations=iterations, factor_scalar=True, trace_cap=iterations)
    wall = time.perf_counter() - t
    tr = st["trace"]
    conv = bool(st["rinf"] <= 1e-6 or st["rinf"] <= 1e-6 * np.abs(u).max())
    diverged = bool(np.nanmax(tr[:, 0]) > 1e3 * r0) or not np.all(np.isfinite(tr[:, 0]))
    status = "converged" if conv else ("diverged" if diverged else "no convergence")
    return {
        "workload": which, "opts": opts, "pressure_pa": m.pressure, "dof": m.ndof,
        "r0_inf": r0, "iterations": st["iterations"], "r_evals": st["r_evals"], "status": status,
        "final_rinf": st["rinf"], "max_rinf": float(np.nanmax(tr[:, 0])), "wall_s": round(wall, 1),
        "factor_s": round(st["factor_seconds"], 2),
        # the first 200 iterations in full, the last 20 as well
        "trace": [[float(a), float(b), int(c)] for a, b, c in tr[:200]],
        "trace_tail": [[float(a), float(b), int(c)] for a, b, c in tr[-20:]],
    }


def main():
    out = {"generator": "tests/golden/make_configured_load_traces.py",
           "oracle": "oracle/efem_oracle.c, projection line search, m = 6, max_linesearch 20, "
                     "epsilon = epsilon_rel = 1e-6, scalar-Laplacian factor",
           "cases": {}}
    for name, which, opts, its in CASES:
        rec = run(which, opts, its)
        print(name, rec["status"], rec["iterations"], "it", rec["wall_s"], "s", flush=True)
        out["cases"][name] = rec
    path = Path(__file__).with_name("configured_load_traces.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
