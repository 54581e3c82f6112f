"""VTK output (Mesh::SaveVTK, Mesh.cpp:292-351, via Cluster::printVTK,
Cluster.cpp:1648-1678): the file written by the native library for a field u is
parsed back and compared with the oracle's mesh (points, hex8 connectivity in the
node order of Mesh.cpp:247-263, partition ids, material ids = remainder(id, 2)),
including uneven slabs and several clusters; the section layout follows the
reference's files."""
import numpy as np
import pytest

import hfls_amd
from oracle.mesh import Mesh, opts

CASES = {
    "cube8_5e3": dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0"),
    "clusters_uneven": dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slab_choices="2,3,4", seed="5", bc="clamp_x0"),
}


def parse(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[2] == "ASCII" and lines[4] == "DATASET UNSTRUCTURED_GRID"
    i = 5
    sec = {}
    while i < len(lines):
        t = lines[i].split()
        if not t:
            i += 1
            continue
        if t[0] == "POINTS":
            n = int(t[1])
            sec["points"] = np.array([[float(v) for v in lines[i + 1 + j].split()] for j in range(n)])
            i += n + 1
        elif t[0] == "CELLS":
            n = int(t[1])
            assert int(t[2]) == 9 * n
            rows = [[int(v) for v in lines[i + 1 + j].split()] for j in range(n)]
            assert all(r[0] == 8 for r in rows)
            sec["cells"] = np.array([r[1:] for r in rows])
            i += n + 1
        elif t[0] == "CELL_TYPES":
            n = int(t[1])
            sec["types"] = np.array([int(lines[i + 1 + j]) for j in range(n)])
            i += n + 1
        elif t[0] == "POINT_DATA":
            n = int(t[1])
            assert lines[i + 1] == "SCALARS displacements float 3" and lines[i + 2] == "LOOKUP_TABLE my_table"
            sec["u"] = np.array([[float(v) for v in lines[i + 3 + j].split()] for j in range(n)])
            i += n + 3
        elif t[0] == "CELL_DATA":
            n = int(t[1])
            assert lines[i + 1] == "SCALARS decomposition int 1" and lines[i + 2] == "LOOKUP_TABLE decomposition"
            sec["part"] = np.array([int(lines[i + 3 + j]) for j in range(n)])
            j0 = i + 3 + n
            assert lines[j0] == "SCALARS materialId int 1" and lines[j0 + 1] == "LOOKUP_TABLE materialId"
            sec["mat"] = np.array([int(lines[j0 + 2 + j]) for j in range(n)])
            i = j0 + 2 + n
        else:
            raise AssertionError("unexpected line %r" % lines[i])
    return sec


@pytest.mark.parametrize("name", sorted(CASES))
def test_vtk_matches_oracle_mesh(name, tmp_path):
    kw = CASES[name]
    p = hfls_amd.Problem(**kw)
    m = Mesh(opts(**kw))
    u = np.sin(np.arange(3 * p.nPoints) * 0.37) * 1e-3
    p.write_vtk(u, tmp_path)
    p.write_vtk(u, tmp_path, iter=3)
    assert (tmp_path / "box3.vtk").read_text() == (tmp_path / "box.vtk").read_text()
    sec = parse(tmp_path / "box.vtk")
    assert np.max(np.abs(sec["points"] - m.points)) <= 5e-7  # "%f": 6 decimals
    assert np.array_equal(sec["cells"], m.elements)
    assert np.all(sec["types"] == 12)
    assert np.max(np.abs(sec["u"] - u.reshape(-1, 3))) <= 5e-7
    assert np.array_equal(sec["part"], m.part)
    assert np.array_equal(sec["mat"], m.material_id[m.part])
    # the reference's "%d " + "% d" cell lines: two spaces after the count
    cells_line = (tmp_path / "box.vtk").read_text().split("CELLS")[1].split("\n")[1]
    assert cells_line.startswith("8  0 1 ")


def test_vtk_rejects_wrong_size(tmp_path):
    p = hfls_amd.Problem(**CASES["cube8_5e3"])
    with pytest.raises(ValueError):
        p.write_vtk(np.zeros(5), tmp_path)
