"""Reference input formats (efem_input_*): an EFEM run directory read into a
problem.  Pinned against an independent Python reading of the same files
(the reference's rules: MeshParser sections, 1-based ids, Body element kinds
with an optional material column, IO::readFile skipping lines with '#' or
'!', faces deduplicated) on the reference's own sample (tests/golden/
sample_beam, copied from sample/beam) and on small synthetic decks for the
edge cases.  No GPU."""
import json
from pathlib import Path

import numpy as np
import pytest

import efem_amd as E

SAMPLE = Path(__file__).resolve().parent / "golden" / "sample_beam"


def py_mesh(path):
    X, el = [], []
    sec = None
    for raw in Path(path).read_text().splitlines():
        l = raw.strip()
        if not l or l.startswith("##") or l.startswith("!!"):
            continue
        if l.startswith("#"):
            sec = "node" if l.startswith("#node") else ("element" if l.startswith("#element") else None)
            continue
        if sec == "node":
            X.append([float(v) for v in l.replace(",", " ").split()[:3]])
        elif sec == "element":
            el.append([int(v) for v in l.replace(",", " ").split()])
    return np.array(X), el


def py_rows(path):
    out = []
    for raw in Path(path).read_text().splitlines():
        l = raw.strip()
        if l and "#" not in l and "!" not in l:
            out.append([int(v) for v in l.replace(",", " ").split()])
    return out


@pytest.mark.parametrize("kind", ["c3d4", "c3d20"])
def test_sample_beam_matches_python_reading(kind, tmp_path):
    m, s = E.Mesh.from_input(SAMPLE / "beam.json") if kind == "c3d4" else _load_abs(tmp_path, kind)
    X, el = py_mesh(SAMPLE / kind / "mesh.txt")
    nen = {4: 4, 8: 8, 10: 10, 20: 20}[len(el[0]) if len(el[0]) in (4, 8, 10, 20) else len(el[0]) - 1]
    assert m.nen == nen and m.n_nodes == len(X) and m.n_elems == len(el)
    assert np.array_equal(m.X, X)
    assert np.array_equal(m.conn, np.array([r[:nen] for r in el]) - 1)
    fixed = {v - 1 for r in py_rows(SAMPLE / kind / "fix.txt") for v in r}
    assert set(np.flatnonzero(m.fixed)) == fixed
    faces = []
    for r in py_rows(SAMPLE / kind / "pressure.txt"):
        f = []
        for v in r:
            if v - 1 not in f:
                f.append(v - 1)
        faces.append(f)
    assert np.array_equal(m.faces, np.array(faces))
    # material id -1 is the mo_riv entry; value [0, 4500, 0] is a directed traction
    assert m.law == E.MOONEY_RIVLIN and np.allclose(m.params, [57000, 57000, 0, 1e-6])
    assert m.pdir is not None and np.allclose(m.pdir, [0, 4500, 0])
    assert s["epsilon"] == 1e-6 and s["max_linesearch"] == 20 and s["use_precond"] and s["m"] == 6


def _load_abs(tmp_path, kind):
    # a copy of the run directory with the c3d20 files
    d = tmp_path / "run"
    (d / kind).mkdir(parents=True)
    for f in ("mesh.txt", "fix.txt", "pressure.txt"):
        (d / kind / f).write_text((SAMPLE / kind / f).read_text())
    cfg = json.loads((SAMPLE / "beam.json").read_text())
    cfg["model"][0]["mesh_file"] = f"{kind}/mesh.txt"
    for c in cfg["constraint"]:
        c["constraint_file"] = f"{kind}/{Path(c['constraint_file']).name}"
    (d / "beam.json").write_text(json.dumps(cfg))
    return E.Mesh.from_input(d / "beam.json")


def test_sample_beam_builds_and_factors_on_host():
    # the problem is accepted by the host setup (every pressure face finds
    # its parent element) and its approximate Jacobian factors
    m, _ = E.Mesh.from_input(SAMPLE / "beam.json")
    b = np.ones(m.ndof)
    x, info = E.host_factor_solve(m, b)
    assert np.all(np.isfinite(x)) and info["nnz"] > 0


def write_deck(d, mesh, fix=None, pressure=None, materials=None, constraints=None, analysis=None):
    d.mkdir(parents=True, exist_ok=True)
    (d / "mesh.txt").write_text(mesh)
    cons = constraints if constraints is not None else []
    if fix is not None:
        (d / "fix.txt").write_text(fix)
        cons.append({"type": "fix", "constraint_file": "fix.txt"})
    if pressure is not None:
        (d / "p.txt").write_text(pressure[0])
        cons.append({"type": "pressure", "value": pressure[1], "constraint_file": "p.txt"})
    cfg = {"model": [{"type": "soft", "mesh_file": "mesh.txt"}], "constraint": cons,
           "material": materials if materials is not None else
           [{"id": -1, "type": "neo_hookean", "C10": 1e5, "D1": 1e-6}]}
    if analysis:
        cfg["analysis"] = analysis
    (d / "run.json").write_text(json.dumps(cfg))
    return d / "run.json"


TWO_TETS = """## two tets sharing a face
#node
0 0 0
1 0 0
0 1 0
0 0 1
1, 1, 1
!! comment inside a section
#element
1, 2, 3, 4
2 3 4 5
"""


def test_edge_cases(tmp_path):
    jp = write_deck(tmp_path / "a", TWO_TETS, fix="# fixed nodes\n1,2\n3 ! third\n",
                    pressure=("2 3 5 3\n", 700.0),
                    analysis={"nl_solver": "lbfgs", "max_solver_iters": 77, "residual_criterion": 1e-8})
    m, s = E.Mesh.from_input(jp)
    assert m.n_nodes == 5 and m.n_elems == 2 and np.allclose(m.X[4], [1, 1, 1])
    assert list(np.flatnonzero(m.fixed)) == [0, 1]      # the '!' line is skipped as a whole
    assert m.faces.tolist() == [[1, 2, 4]]              # repeated node dropped
    assert m.pressure == 700.0 and m.pdir is None       # scalar value: follower pressure
    assert not s["use_precond"] and s["max_iterations"] == 77 and s["epsilon"] == 1e-8


def test_material_column_and_promotion(tmp_path):
    mesh = TWO_TETS.replace("1, 2, 3, 4", "1, 2, 3, 4, 7").replace("2 3 4 5", "2 3 4 5 9")
    mats = [{"id": 7, "type": "neo_hookean", "C10": 1e5, "D1": 1e-6},
            {"id": 9, "type": "mo_riv", "C10": 6e4, "C01": 2e4, "D1": 2e-6}]
    m, _ = E.Mesh.from_input(write_deck(tmp_path / "b", mesh, materials=mats))
    assert m.law == E.MOONEY_RIVLIN and m.params.shape == (2, 4)
    assert np.allclose(m.params, [[1e5, 0, 0, 1e-6], [6e4, 2e4, 0, 2e-6]])
    # an element whose material is not defined is skipped (as the reference)
    m2, _ = E.Mesh.from_input(write_deck(tmp_path / "c", mesh, materials=mats[:1]))
    assert m2.n_elems == 1 and m2.skipped_elements == 1


@pytest.mark.parametrize("what", ["mr_yeoh", "missing_mesh", "bad_json", "bad_node"])
def test_refused_inputs(what, tmp_path):
    d = tmp_path / what
    if what == "mr_yeoh":
        mesh = TWO_TETS.replace("1, 2, 3, 4", "1, 2, 3, 4, 1").replace("2 3 4 5", "2 3 4 5 2")
        jp = write_deck(d, mesh, materials=[{"id": 1, "type": "mo_riv", "C10": 1, "C01": 1, "D1": 1},
                                            {"id": 2, "type": "yeoh", "C10": 1, "C20": 1, "C30": 1, "D1": 1}])
    elif what == "missing_mesh":
        jp = write_deck(d, TWO_TETS)
        (d / "mesh.txt").unlink()
    elif what == "bad_json":
        jp = write_deck(d, TWO_TETS)
        jp.write_text('{"model": [ {"mesh_file": "mesh.txt" }')
    else:
        jp = write_deck(d, TWO_TETS.replace("1, 1, 1", "1, x, 1"))
    with pytest.raises(E.EfemError):
        E.Mesh.from_input(jp)
