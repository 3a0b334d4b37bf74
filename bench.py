#!/usr/bin/env python3
"""Headline benchmark: time-to-solution and residual evaluations per second
of the L-BFGS solve loop on the 265,923-DOF beam (200x20x20 hexes split into
480,000 Kuhn C3D4 tets, interior nodes jittered +-10 %, Mooney-Rivlin
C10=80000 C01=20000 D1=1e-6, x=0 fixed, follower pressure on z=max), mixed
precision vs FP64, on one MI355X.

A step is one complete solve from u = 0: approximate-Jacobian assembly,
ordering and factorisation (host) + upload + the L-BFGS loop on the device,
stopped when the displacement is within 1e-6 (relative L2) of a tightly
converged FP64 solution u*.  The reference's own stopping test is kept too.

Load level: at the configured 3 kPa the reference algorithm (projection line
search, m = 6, approximate-Jacobian H0) does not converge -- the residual
grows without bound on this mesh, and on coarser meshes of the same beam
even with 20 load increments -- so the time-to-solution is measured at the
largest follower pressure it solves in one step (--pressure, default
100 Pa) on the same mesh; throughput does not depend on the load level.

Prints ONE JSON line (see the field list at the end of main()).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
P_CONFIGURED = 3000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def element_work(m, prec):
    """Algorithmic bytes / flops of one residual evaluation (C3D4, MR):
    B = 3 N_f s_u + 3 N s_X + 4 n_e E + s E (V_e) + 8 * 3 N_f (r)
    F = E g (72 n_e + 370)."""
    s = 4 if prec == "mixed" else 8
    nf = m.n_free
    B = 3 * nf * s + 3 * m.n_nodes * s + 4 * m.nen * m.n_elems + s * m.n_elems + 24 * nf
    F = m.n_elems * 1 * (72 * m.nen + 370)
    return B, F


def cpu_baseline(mesh, n_evals):
    """The oracle (single-thread C restatement of the reference path) on the
    same mesh: residual evaluations per second on `n_evals` evaluations."""
    import oracle as O
    om = O.Model(mesh)
    rng = np.random.default_rng(1)
    u = 1e-5 * rng.standard_normal(om.ndof)
    om.residual(u)
    t0 = time.perf_counter()
    for _ in range(n_evals):
        om.residual(u)
    dt = time.perf_counter() - t0
    return n_evals / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pressure", type=float, default=100.0)
    ap.add_argument("--target", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--cpu-evals", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-switch", type=float, default=1e-3,
                    help="mixed: element kernels go FP64 once |r|inf <= this x |r0|inf")
    ap.add_argument("--profile-run", type=int, default=0,
                    help="only run one mixed solve of this many iterations (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU design "
                                   "(mesh-partition halo exchange) is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import efem_amd as E

    mesh = E.Mesh.workload(2, [200, 20, 20, 20240531, args.pressure])
    log(f"beam: {mesh.n_nodes} nodes, {mesh.n_elems} C3D4, {mesh.ndof} DOF, p = {args.pressure} Pa")
    if args.profile_run:
        c = E.Context(mesh, E.MIXED)
        c.profile_reset()
        t = time.perf_counter()
        u, st = c.solve(epsilon=0.0, epsilon_rel=0.0, max_iterations=args.profile_run)
        dt = time.perf_counter() - t
        ems, ek = c.profile_read()
        hms, hk = c.profile_read_h0()
        print(json.dumps({"profile_run": True, "iterations": st["iterations"], "r_evals": st["r_evals"],
                          "wall_s": dt, "factor_s": st["factor_seconds"], "elem_us": 1e3 * ems / max(ek, 1),
                          "elem_launches": ek, "h0_ms": hms / max(hk, 1), "h0_applies": hk}))
        return
    ctx = {"mixed": E.Context(mesh, E.MIXED)}
    if not args.no_fp64:
        ctx["fp64"] = E.Context(mesh, E.FP64)
    c64 = ctx.get("fp64") or E.Context(mesh, E.FP64)

    # u*: FP64, converged far below the 1e-6 displacement target
    fext = np.abs(c64.residual(np.zeros(mesh.ndof))).max()
    t0 = time.perf_counter()
    u_star, st_star = c64.solve(epsilon=1e-10 * fext, epsilon_rel=0.0, max_iterations=args.max_iters)
    log(f"u*: {st_star['iterations']} it, |r|inf {st_star['rinf']:.3e} "
        f"(f_ext {fext:.3e}), {time.perf_counter() - t0:.1f}s")

    def one_solve(name, profile=False):
        c = ctx[name]
        c.factor_release()
        if profile:
            c.profile_reset()
        t = time.perf_counter()
        u, st = c.solve(epsilon=0.0, epsilon_rel=0.0, max_iterations=args.max_iters,
                        x_target=u_star, target_tol=args.target,
                        fp64_switch=args.fp64_switch if name == "mixed" else 0.0)
        dt = time.perf_counter() - t
        st["wall"] = dt
        if not st["target_reached"]:
            log(f"WARNING {name}: target not reached (err {st['target_err']:.2e})")
        return st

    for _ in range(args.warmup):
        for name in ctx:
            one_solve(name)

    res = {name: [] for name in ctx}
    prof = {}
    for k in range(args.steps):
        for name in ctx:
            st = one_solve(name, profile=(k == args.steps - 1))
            res[name].append(st)
            log(f"{name}: {st['iterations']} it, {st['r_evals']} evals, err {st['target_err']:.2e}, "
                f"factor {st['factor_seconds']:.2f}s, total {st['wall']:.3f}s, "
                f"fp64 from it {st['fp64_switch_iteration']}")
            if k == args.steps - 1:
                ems, ek = ctx[name].profile_read()
                hms, hk = ctx[name].profile_read_h0()
                prof[name] = (ems, ek, hms, hk, ctx[name].factor_nnz)

    mix = res["mixed"]
    t_mix = float(np.mean([s["wall"] for s in mix]))
    evals_mix = sum(s["r_evals"] for s in mix)
    loop_mix = sum(s["wall"] - s["factor_seconds"] for s in mix)
    value = evals_mix / sum(s["wall"] for s in mix)
    out = {
        "metric": "time-to-solution & residual evals/s: mixed vs FP64 on GPU at 1e-6 disp err",
        "value": round(value, 2),
        "unit": "residual evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 (mixed)",
        "data": "synthetic",
        "config": {
            "workload": "beam 0.2x0.02x0.02 m, 200x20x20 hexes -> 480000 Kuhn C3D4 (88641 nodes, "
                        "265923 DOF), jitter +-10%, Mooney-Rivlin 80000/20000/1e-6, x=0 fixed, "
                        "follower pressure on z=max",
            "pressure_pa": args.pressure,
            "pressure_configured_pa": P_CONFIGURED,
            "pressure_note": "reference L-BFGS diverges at the configured load; see DESIGN.md",
            "target_disp_err": args.target,
            "lbfgs": {"m": 6, "max_linesearch": 20, "line_search": "projection (reference)"},
            "parallelism": "single GPU",
        },
        "tts_mixed_s": round(t_mix, 4),
        "loop_evals_per_s_mixed": round(evals_mix / loop_mix, 2),
        "iterations_mixed": int(np.mean([s["iterations"] for s in mix])),
        "r_evals_mixed": int(np.mean([s["r_evals"] for s in mix])),
        "factor_s_mixed": round(float(np.mean([s["factor_seconds"] for s in mix])), 3),
        "disp_err_mixed": float(max(s["target_err"] for s in mix)),
        "fp64_switch": args.fp64_switch,
        "fp64_switch_iteration": int(mix[-1]["fp64_switch_iteration"]),
        "u_star_iterations": st_star["iterations"],
    }
    if "fp64" in res:
        f64 = res["fp64"]
        t64 = float(np.mean([s["wall"] for s in f64]))
        out.update({
            "tts_fp64_s": round(t64, 4),
            "speedup_mixed_vs_fp64": round(t64 / t_mix, 3),
            "evals_per_s_fp64": round(sum(s["r_evals"] for s in f64) / sum(s["wall"] for s in f64), 2),
            "iterations_fp64": int(np.mean([s["iterations"] for s in f64])),
            "disp_err_fp64": float(max(s["target_err"] for s in f64)),
        })
    # roofline of the dominant kernel family: the preconditioner apply
    # (forward + backward supernodal triangular solves), HBM-bound
    ems, ek, hms, hk, nnz = prof["mixed"]
    h0_ms = hms / max(hk, 1)
    h0_bytes = 2 * nnz * 4 + 48 * mesh.n_free
    out["roofline"] = {
        "kernel": "H0 apply (kTrsvFwd + kTrsvBwd levels), mixed",
        "bound": "hbm",
        "achieved": round(h0_bytes / (h0_ms * 1e-3) / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(h0_bytes / (h0_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": h0_bytes,
        "avg_ms": round(h0_ms, 4),
        "factor_nnz": nnz,
    }
    B, F = element_work(mesh, "mixed")
    e_us = 1e3 * ems / max(ek, 1)
    t_roof = max(B / (HBM_PEAK_GBS * 1e9), F / (FP32_PEAK_TFLOPS * 1e12))
    out["roofline_element"] = {
        "kernel": "kElemTet4<float, MR>",
        "bound": "hbm" if B / HBM_PEAK_GBS > F / (FP32_PEAK_TFLOPS * 1e3) else "valu",
        "achieved_GBs": round(B / (e_us * 1e-6) / 1e9, 1),
        "achieved_TFLOPs": round(F / (e_us * 1e-6) / 1e12, 2),
        "frac": round(t_roof / (e_us * 1e-6), 4),
        "avg_us": round(e_us, 2),
        "alg_bytes": B,
        "alg_flops": F,
    }
    if "fp64" in prof:
        ems64, ek64, hms64, hk64, _ = prof["fp64"]
        B64, F64 = element_work(mesh, "fp64")
        e64 = 1e3 * ems64 / max(ek64, 1)
        out["roofline_element_fp64"] = {
            "avg_us": round(e64, 2),
            "achieved_TFLOPs": round(F64 / (e64 * 1e-6) / 1e12, 2),
            "frac": round(max(B64 / (HBM_PEAK_GBS * 1e9), F64 / (FP64_PEAK_TFLOPS * 1e12)) / (e64 * 1e-6), 4),
        }
        out["h0_apply_ms_fp64"] = round(hms64 / max(hk64, 1), 4)
    if not args.no_cpu:
        rate, dt = cpu_baseline(mesh, args.cpu_evals)
        out["cpu_baseline"] = {
            "value": round(rate, 3),
            "unit": "residual evals/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_evals} residual evaluations of the same 480000-element mesh by the "
                      f"single-thread C oracle ({dt:.1f} s); host {platform.processor() or platform.machine()}, "
                      f"{os.cpu_count()} logical cores visible",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
