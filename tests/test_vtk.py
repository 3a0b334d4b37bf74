"""Legacy-VTK output (Model::saveVTK, src/model/Model.cpp:50-117) and the
results-file naming, host only.  The file is parsed back with an
independent reader of the legacy format: the points must be the FP32
values of X + u at every node (fixed nodes at X), the cells the element
node lists unchanged, the cell types VTK's codes for the element kind."""
import datetime as dt
import re

import numpy as np
import pytest

import efem_amd as E

VTK_TYPE = {4: 10, 10: 24, 8: 12, 20: 25}


def read_legacy_vtk(path):
    tok = open(path).read().split("\n")
    assert tok[0].startswith("# vtk DataFile Version")
    assert tok[2].strip() == "ASCII" and tok[3].strip() == "DATASET UNSTRUCTURED_GRID"
    words = " ".join(tok[4:]).split()
    i = 0
    out = {}
    while i < len(words):
        key = words[i]
        if key == "POINTS":
            n, dtype = int(words[i + 1]), words[i + 2]
            assert dtype == "float"
            out["points"] = np.array(words[i + 3: i + 3 + 3 * n], np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif key == "CELLS":
            nc, size = int(words[i + 1]), int(words[i + 2])
            flat = np.array(words[i + 3: i + 3 + size], np.int64)
            cells, k = [], 0
            for _ in range(nc):
                cells.append(flat[k + 1: k + 1 + flat[k]].tolist())
                k += 1 + flat[k]
            assert k == size
            out["cells"] = cells
            i += 3 + size
        elif key == "CELL_TYPES":
            nc = int(words[i + 1])
            out["types"] = np.array(words[i + 2: i + 2 + nc], np.int64)
            i += 2 + nc
        else:
            raise AssertionError(f"unexpected section {key}")
    return out


def scatter(mesh, u):
    """node displacements from a free-DOF vector (free ids in node order)"""
    fid = E.free_ids(mesh.fixed)
    d = np.zeros((mesh.n_nodes, 3))
    free = fid >= 0
    d[free] = np.asarray(u).reshape(-1, 3)[fid[free]]
    return d


@pytest.mark.parametrize("kind,law", [(4, 0), (10, 0), (8, 1), (20, 2)])
def test_deformed_grid_round_trip(kind, law, tmp_path, rng):
    m = E.Mesh.workload(0, [3, 2, 2, 0.003, 0.002, 0.002, kind, law, 1.0, 0.0, 5, 0])
    u = rng.standard_normal(m.ndof) * 1e-4
    p = tmp_path / "f.vtk"
    E.write_vtk(m, u, p)
    g = read_legacy_vtk(p)
    # 9 significant digits: read back as FP32, the text is the same value
    want = (m.X + scatter(m, u)).astype(np.float32)
    assert np.array_equal(g["points"].astype(np.float32), want)
    assert g["cells"] == m.conn.tolist()
    assert np.all(g["types"] == VTK_TYPE[kind])
    # fixed nodes stay where they are
    fixed = m.fixed.astype(bool)
    assert fixed.any() and np.array_equal(g["points"][fixed].astype(np.float32), m.X[fixed].astype(np.float32))


def test_undeformed_frame(tmp_path):
    m = E.Mesh.workload(0, [2, 2, 2, 0.002, 0.002, 0.002, 4, 0, 1.0, 0.0, 5, 0])
    p = tmp_path / "u.vtk"
    E.write_vtk(m, None, p)
    g = read_legacy_vtk(p)
    assert np.array_equal(g["points"].astype(np.float32), m.X.astype(np.float32))
    assert len(g["cells"]) == m.n_elems


def test_sample_deck_frame(tmp_path):
    # the reference's sample run: every node is a point (unused ones too),
    # the element rows as MeshParser read them
    from pathlib import Path
    m, _ = E.Mesh.from_input(Path(__file__).resolve().parent / "golden" / "sample_beam" / "beam.json")
    p = tmp_path / "beam.vtk"
    E.write_vtk(m, np.full(m.ndof, 1e-3), p)
    g = read_legacy_vtk(p)
    assert g["points"].shape == (m.n_nodes, 3) and g["cells"] == m.conn.tolist()
    assert np.all(g["types"] == VTK_TYPE[m.nen])


def test_errors(tmp_path):
    m = E.Mesh.workload(0, [2, 2, 2, 0.002, 0.002, 0.002, 4, 0, 1.0, 0.0, 5, 0])
    with pytest.raises(ValueError):
        E.write_vtk(m, np.zeros(m.ndof + 1), tmp_path / "x.vtk")
    with pytest.raises(E.EfemError, match="cannot open"):
        E.write_vtk(m, None, tmp_path / "missing_dir" / "x.vtk")
    bad = E.Mesh.workload(0, [2, 2, 2, 0.002, 0.002, 0.002, 4, 0, 1.0, 0.0, 5, 0])
    bad.conn = bad.conn.copy()
    bad.conn[0, 0] = bad.n_nodes
    with pytest.raises(E.EfemError, match="out of range"):
        E.write_vtk(bad, None, tmp_path / "y.vtk")


def test_results_path(tmp_path):
    before = dt.datetime.now()
    p = E.results_path(str(tmp_path) + "/", "deformed")
    after = dt.datetime.now()
    assert p.parent.is_dir() and p.parent.parent == tmp_path / "results"
    day = p.parent.name
    assert re.fullmatch(r"\d{4}-\d{2}-\d{2}", day)
    mt = re.fullmatch(r"(\d{4}-\d{2}-\d{2})-(\d{2})-(\d{2})-(\d{2})-(\d{3})-deformed\.vtk", p.name)
    assert mt and mt.group(1) == day
    t = dt.datetime.strptime(p.name[:23], "%Y-%m-%d-%H-%M-%S-%f")
    assert before.replace(microsecond=before.microsecond // 1000 * 1000) - dt.timedelta(seconds=1) <= t <= after
    # a working directory without the trailing separator names the same place
    q = E.results_path(str(tmp_path), "undeformed")
    assert q.parent.parent == tmp_path / "results" and q.name.endswith("-undeformed.vtk")
