"""The reference's problem files (fixtures under tests/golden/problems, the
reference's tests/problems unchanged) and Gmsh meshes (tests/golden/meshes):
every file loads into this build's objects or is refused as outside the
operator path (1D, CG-SUPG, the Dirichlet-cell vacuum treatment), the schema's
error cases are errors, and the Gmsh reader and the painting give what the
files say."""
import glob
import json
import os

import numpy as np
import pytest

from ubolt_amd import problem_file as pfm

HERE = os.path.dirname(os.path.abspath(__file__))
PROBLEMS = os.path.join(HERE, "golden", "problems")
MESHES = os.path.join(HERE, "golden", "meshes")
FILES = sorted(glob.glob(os.path.join(PROBLEMS, "*.json")))


def expected_refusal(spec):
    mesh = spec["mesh"]
    if spec["dimension"] == 1:
        return "1D"
    if mesh.get("discretisation") == "cg_supg":
        return "CG-SUPG"
    if spec.get("vacuum_treatment") == "dirichlet_cell":
        return "Dirichlet-cell"
    return None


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_every_reference_problem_file(path):
    spec = json.load(open(path))
    why = expected_refusal(spec)
    if why:
        with pytest.raises(pfm.UnsupportedProblem, match=why):
            pfm.load(path)
        return
    pf = pfm.load(path)
    p = pf.problem
    d = p.disc
    assert pf.dim == spec["dimension"] and d.dim == spec["dimension"]
    assert p.n_groups == spec["materials"]["n_groups"] if isinstance(spec["materials"], dict) else True
    assert p.sigma_t.shape == (p.n_groups, d.n_cells) and p.sigma_s.shape == (p.n_groups, p.n_groups, d.n_cells)
    if "n_cells" in spec["mesh"]:
        n = int(np.prod(spec["mesh"]["n_cells"]))
        factor = 1 if not spec["mesh"].get("simplex") else (2 if spec["dimension"] == 2 else 6)
        assert d.n_cells == n * factor
    assert d.nb == (1 if spec["mesh"].get("order", 0) == 0 else spec["dimension"] + 1)
    assert np.all(p.sigma_t >= 0)


def test_refusal_count():
    """of the reference's 121 problem files, 79 are on this build's path"""
    ok = sum(expected_refusal(json.load(open(f))) is None for f in FILES)
    assert len(FILES) == 121 and ok == 79


def test_gmsh_square_2x2():
    m = pfm.read_gmsh(os.path.join(MESHES, "square_2x2_tri.msh"))
    cells, phys = m.cells(2)
    assert m.nodes.shape == (9, 3) and cells.shape == (8, 3)
    assert sorted(set(phys.tolist())) == [1, 2]
    lines = m.faces(2)
    tags = np.concatenate([p for _, p in lines])
    assert sorted(tags.tolist()) == [10, 10, 11, 11, 12, 12, 13, 13]
    assert m.physical_names[(1, 13)] == "left" and m.physical_names[(2, 2)] == "right_half"
    p = m.nodes[cells][:, :, :2]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0]))
    assert abs(area.sum() - 1.0) < 1e-15
    # "left_half" is x < 0.5: every vertex of a set-1 cell has x <= 0.5
    assert np.all(m.nodes[cells[phys == 1]][:, :, 0] <= 0.5) and np.all(m.nodes[cells[phys == 2]][:, :, 0] >= 0.5)


@pytest.mark.parametrize("name", ["square_irregular_tri.msh", "tri_slanted.msh"])
def test_gmsh_meshes_read(name):
    m = pfm.read_gmsh(os.path.join(MESHES, name))
    cells, _ = m.cells(2)
    p = m.nodes[cells][:, :, :2]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0]))
    assert np.all(area > 0) and len(m.faces(2)) >= 1


def test_file_mesh_problem_cell_sets():
    pf = pfm.load(os.path.join(PROBLEMS, "plex_square_msh_dg1.json"))
    spec = json.load(open(os.path.join(PROBLEMS, "plex_square_msh_dg1.json")))
    names = {m["name"]: m["id"] for m in spec["materials"]["materials"]}
    want = np.where(pf.cell_sets == 1, names["scatterer"], names["absorber"])
    assert np.array_equal(pf.material, want)
    assert pf.problem.disc.n_cells == 8 and pf.problem.disc.nb == 3


def test_paint_by_centre():
    """box_50_absorber: the absorber box [0.4, 0.6]^2 over a 50 x 50 grid holds
    the cells whose centres (i + 0.5) / 50 lie in it: i = 20..29 on each axis"""
    pf = pfm.load(os.path.join(PROBLEMS, "box_50_absorber.json"))
    m = pf.material.reshape(50, 50)
    want = np.zeros((50, 50), dtype=int)
    want[20:30, 20:30] = 1
    assert np.array_equal(m, want)
    assert np.all(pf.problem.sigma_t[0].reshape(50, 50)[want == 1] == 10.0)


def test_materials_by_path_and_upscatter():
    pf = pfm.load(os.path.join(PROBLEMS, "box_up3_layers.json"))
    assert pf.problem.n_groups == 3 and pf.upscatter


def _write(tmp_path, spec):
    p = tmp_path / "p.json"
    p.write_text(json.dumps(spec))
    return str(p)


BASE = {"dimension": 2, "mesh": {"n_cells": [4, 4], "lengths": [1.0, 1.0]}, "sn_order": 2,
        "materials": {"n_groups": 1, "materials": [{"id": 0, "Sigma_t": [1.0], "Sigma_s": [[0.5]], "Source": [1.0]}]}}


@pytest.mark.parametrize("change,msg", [
    (lambda s: s.update(extra=1), "unknown key"),
    (lambda s: s["mesh"].update(nx=4), "unknown key"),
    (lambda s: s.update(inflow=1.0), "top-level 'inflow'"),
    (lambda s: s.update(boundary_conditions={"left": {"type": "vacuum", "window": [0.0, 0.5]}}), "without an inflow"),
    (lambda s: s.update(boundary_conditions={"left": {"type": "reflect", "inflow": 1.0}}), "no inflow or window"),
    (lambda s: s.update(boundary_conditions={"left": {"type": "vacuum", "inflow": 1.0, "window": [0.6, 0.5]}}),
     "pairs"),
    (lambda s: s.update(boundary_conditions={"4": "reflect"}), "unstructured meshes only"),
    (lambda s: s.update(boundary_conditions={"west": "reflect"}), "unknown face"),
    (lambda s: s.update(sn_order=3), "positive even"),
    (lambda s: s["mesh"].update(order=1), "unstructured meshes only"),
    (lambda s: s.update(regions={"paint": [{"material": "nope", "box": [0, 1, 0, 1]}]}), "no material"),
    (lambda s: s.update(regions={"cell_sets": {"1": 0}}), "unstructured meshes only"),
])
def test_schema_errors(tmp_path, change, msg):
    spec = json.loads(json.dumps(BASE))
    change(spec)
    with pytest.raises(pfm.ProblemError, match=msg):
        pfm.load(_write(tmp_path, spec))


def test_unstructured_twin_face_ids(tmp_path):
    """a name and an integer reaching the same 'Face Sets' id is an error"""
    spec = json.loads(json.dumps(BASE))
    spec["mesh"]["type"] = "unstructured"
    spec["boundary_conditions"] = {"left": "reflect", "4": "vacuum"}
    with pytest.raises(pfm.ProblemError, match="second time"):
        pfm.load(_write(tmp_path, spec))
