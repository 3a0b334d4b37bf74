"""GPU: batches of independent instances (efem_batch_*).

Pins:
  * the merged residual equals the oracle's residual of every instance
    (<= 1e-10, FP64) -- the concatenation adds nothing;
  * the merged preconditioner apply equals each instance's own context
    bit for bit (same per-instance factor, same work items);
  * the persistent batched L-BFGS: an instance's result is bitwise the same
    whatever else is in the batch (fixed per-workgroup reduction trees);
    it reaches the converged solution of the grid-path solve of that
    instance alone (<= 1e-6) with a comparable iteration count; mixed
    precision reaches the FP64 batch solution.
"""
import numpy as np
import pytest

import efem_amd as E
import oracle as O
from helpers import rel

pytestmark = pytest.mark.gpu


def beams(idx, seed=0, shorten=True):
    """Batch instances (workload 5); shortened for the parity tests."""
    out = []
    for i in idx:
        m = E.Mesh.workload(5, [i, seed])
        if shorten:
            # keep at most 12 elements along x: cheap oracle solves
            nx = m.n_elems // 16
            if nx > 12:
                keep = m.X[:, 0] <= 12 * 0.005 + 1e-9
                newid = -np.ones(m.n_nodes, np.int64)
                newid[keep] = np.arange(keep.sum())
                ek = np.all(keep[m.conn], axis=1)
                fk = np.all(keep[m.faces], axis=1)
                m = E.Mesh(m.X[keep], newid[m.conn[ek]], m.law, m.params, m.fixed[keep], newid[m.faces[fk]],
                           m.pressure, m.reduced)
        out.append(m)
    return out


def test_batch_residual_matches_oracle(rng):
    ms = beams(range(5))
    with E.Batch(ms) as b:
        u = 1e-4 * rng.standard_normal(b.ndof)
        r = b.split(b.residual(u))
        for i, (m, ui) in enumerate(zip(ms, b.split(u))):
            assert rel(r[i], O.Model(m).residual(ui)) <= 1e-10, i


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_batch_apply_h0_matches_single(prec, rng):
    ms = beams(range(4))
    with E.Batch(ms, prec) as b:
        assert b.factor()
        v = rng.standard_normal(b.ndof)
        hv = b.split(b.apply_h0(v))
        for i, (m, vi) in enumerate(zip(ms, b.split(v))):
            with E.Context(m, prec) as c:
                # a batch factors its instances on the host and sweeps
                # every level, root included
                c.set_option("device_factor", 0)
                c.set_option("fused_root", 0)
                assert c.factor()
                assert np.array_equal(hv[i], c.apply_h0(vi)), i


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_batch_result_independent_of_batch(prec):
    ms = beams(range(6))
    kw = dict(epsilon=1e-7, max_iterations=3000, fp64_switch=1e-3 if prec == E.MIXED else 0.0)
    with E.Batch(ms, prec) as b:
        u, st = b.solve(**kw)
        parts = b.split(u)
    with E.Batch([ms[3]], prec) as b1:
        u3, st3 = b1.solve(**kw)
    assert np.array_equal(u3, parts[3])
    assert st3[0]["iterations"] == st[3]["iterations"] and st3[0]["r_evals"] == st[3]["r_evals"]
    with E.Batch([ms[5], ms[3]], prec) as b2:
        u53, st53 = b2.solve(**kw)
        assert np.array_equal(b2.split(u53)[1], parts[3])


def test_batch_matches_single_grid_solves():
    ms = beams(range(6))
    with E.Batch(ms) as b:
        u, st = b.solve(epsilon=1e-8, max_iterations=5000)
        parts = b.split(u)
    for i, m in enumerate(ms):
        assert st[i]["rinf"] <= 1e-8, (i, st[i])
        with E.Context(m) as c:
            ui, si = c.solve(epsilon=1e-8, max_iterations=5000)
        print(f"instance {i}: batch {st[i]['iterations']} it, grid {si['iterations']} it", flush=True)
        assert rel(parts[i], ui) <= 1e-6, i
        assert abs(st[i]["iterations"] - si["iterations"]) <= 0.2 * si["iterations"] + 5


def test_batch_matches_oracle_solve():
    ms = beams([1, 2])
    with E.Batch(ms) as b:
        u, st = b.solve(epsilon=1e-8, max_iterations=5000)
        parts = b.split(u)
    for i, m in enumerate(ms):
        uo, so = O.Model(m).solve(epsilon=1e-8, max_iterations=5000)
        assert rel(parts[i], uo) <= 1e-6, i


def test_batch_mixed_reaches_fp64_solution():
    ms = beams(range(4))
    with E.Batch(ms) as b:
        u64, _ = b.solve(epsilon=1e-9, max_iterations=5000)
    with E.Batch(ms, E.MIXED) as b:
        u32, st = b.solve(epsilon=1e-9, max_iterations=5000, fp64_switch=1e-3)
    for a, c in zip(b.split(u32), b.split(u64)):
        assert rel(a, c) <= 1e-6


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_batch_dense_preconditioner_matches_level_sweeps(prec):
    # the dense per-instance Z_i = (L_i L_i^T)^-1 (option batch_dense_h0, the
    # default) against the level-scheduled sweeps: the same K_i^-1 up to
    # rounding, so both solves reach the same converged displacements with
    # comparable iteration counts; full-length beams (675 to 725 free nodes,
    # loads x 1/64 as in the bench); an instance's result does not depend on
    # the rest of the batch
    ms = [E.Mesh.workload(5, [i, 20240531, 1.0 / 64]) for i in (0, 5, 9)]
    out = {}
    for dense in (1, 0):
        with E.Batch(ms, prec) as b:
            b.set_option("batch_dense_h0", dense)
            u, st = b.solve(epsilon=1e-9, max_iterations=8000, fp64_switch=1e-3 if prec == E.MIXED else 0.0)
            out[dense] = (b.split(u), st)
            if dense:
                with E.Batch(ms[1:2], prec) as one:
                    one.set_option("batch_dense_h0", 1)
                    u1, _ = one.solve(epsilon=1e-9, max_iterations=8000,
                                      fp64_switch=1e-3 if prec == E.MIXED else 0.0)
                assert np.array_equal(u1, out[1][0][1]), "dense path not batch-independent"
    for i in range(len(ms)):
        assert out[1][1][i]["rinf"] <= 1e-9 and out[0][1][i]["rinf"] <= 1e-9, i
        assert rel(out[1][0][i], out[0][0][i]) <= 1e-6, i
        it1, it0 = out[1][1][i]["iterations"], out[0][1][i]["iterations"]
        assert abs(it1 - it0) <= 0.3 * max(it0, 1), (i, it1, it0)


def test_batch_rejects_bad_input():
    ms = beams(range(2))
    with E.Batch(ms) as b:
        with pytest.raises(E.EfemError):
            b.solve(max_iterations=0)
        with pytest.raises(E.EfemError):
            E.Context.solve(b)
    with pytest.raises(E.EfemError):
        E.Batch([ms[0], E.Mesh.workload(2, [4, 2, 2])])  # C3D8 with C3D4


@pytest.mark.parametrize("width", [256, 512, 1024])
@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_batch_widths_match_oracle_and_are_batch_independent(width, prec):
    # every workgroup width (wider workgroups split each item's columns and
    # the vector loops over more lanes, so sums are ordered differently) gives
    # the oracle's equilibrium, and an instance's result does not depend on
    # the rest of the batch
    ms = beams([0, 3, 5])
    kw = dict(epsilon=1e-8, max_iterations=5000, fp64_switch=1e-3 if prec == E.MIXED else 0.0)
    print(f"batch width {width} prec {prec}", flush=True)
    with E.Batch(ms, prec) as b:
        b.set_option("batch_threads", width)
        u, st = b.solve(**kw)
        parts = b.split(u)
    for i, m in enumerate(ms):
        assert st[i]["rinf"] <= 1e-8, (i, st[i])
        uo, _ = O.Model(m).solve(epsilon=1e-9, max_iterations=5000)
        assert rel(parts[i], uo) <= 1e-6, i
    with E.Batch([ms[1]], prec) as b1:
        b1.set_option("batch_threads", width)
        u1, st1 = b1.solve(**kw)
    assert np.array_equal(u1, parts[1]) and st1[0]["iterations"] == st[1]["iterations"]


def test_batch_stops_each_instance_at_its_displacement_target():
    # per-instance u* (FP64 grid solves to 1e-10): the batch stops every
    # instance once it is within 1e-6 of its own u*
    ms = beams(range(5))
    us = []
    for m in ms:
        with E.Context(m) as c:
            us.append(c.solve(epsilon=1e-10, max_iterations=5000)[0])
    with E.Batch(ms, E.MIXED) as b:
        u, st = b.solve(epsilon=0.0, epsilon_rel=0.0, max_iterations=5000, fp64_switch=1e-3,
                        x_target=np.concatenate(us), target_tol=1e-6)
        parts = b.split(u)
    for i in range(len(ms)):
        assert st[i]["target_reached"] == 1, (i, st[i])
        err = rel(parts[i], us[i])
        assert err <= 1e-6 and abs(err - st[i]["target_err"]) <= 1e-9, (i, err, st[i]["target_err"])


def test_batch_autotune_picks_a_width():
    ms = beams(range(8))
    with E.Batch(ms, E.MIXED) as b:
        rep = b.autotune()
        assert rep["batch_threads"] in (256, 512, 1024) and rep["batch_ms"] > 0
        u, st = b.solve(epsilon=1e-8, max_iterations=5000, fp64_switch=1e-3)
        assert all(s["rinf"] <= 1e-8 for s in st)


def test_batch_mixed_instances_switching_at_different_times():
    # instances of very different lengths switch from FP32 to FP64 elements
    # at different iterations; their element force slots must not overlap
    # (each precision has its own buffer): every instance equals the same
    # instance solved in a batch of its own, bit for bit
    ms = beams([0, 1, 2, 3, 4, 5, 6, 7], shorten=False)
    kw = dict(epsilon=1e-9, max_iterations=5000, fp64_switch=1e-2)
    with E.Batch(ms, E.MIXED) as b:
        u, st = b.solve(**kw)
        parts = b.split(u)
    sw = sorted(s["fp64_switch_iteration"] for s in st)
    assert sw[0] != sw[-1]
    for i in (0, 3, 7):
        with E.Batch([ms[i]], E.MIXED) as b1:
            ui, si = b1.solve(**kw)
        assert np.array_equal(ui, parts[i]) and si[0]["iterations"] == st[i]["iterations"], i
