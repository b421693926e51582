"""Solution files in the reference's .pos format (simpleMesh::writeSolution /
readSolution, simpleMesh.cc:322-459; PRINTER field names, printer.cc:11-80).

The printed fields come from the oracle's formatter (pinned against the
reference's Lformater in test_oracle_vs_reference.py).  Checked here: the
file names and header lines the reference writes, the element ids of the
mesh, a write/read round trip bit for bit (20 significant digits), and
refusal of files that do not fit the mesh.  No GPU.
"""
import numpy as np
import pytest

import oracle
from dg_amd import (FARFIELD, MULTIFLUID, PASSIVE, PERIODIC, REFLECTIVE, ROE, SINGLEFLUID, STIFFENED, DgxError, Mesh,
                    nfields, problems)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("ph", [SINGLEFLUID, MULTIFLUID, PASSIVE, STIFFENED])
def test_pos_files_round_trip(tmp_path, dim, ph):
    nf = nfields(dim, ph)
    if dim == 1:
        m = Mesh(1, 2, 12, 1, -1.0, 1.0, bc=(FARFIELD, REFLECTIVE), nfields=nf)
    else:
        m = Mesh(2, 2, 6, 4, 0.0, 1.5, 0.0, 1.0, bc=(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=nf)
    t = m.tables()
    out = oracle.Problem(t, ph, ROE).format(problems.riemann_state(m, ph, seed=3))
    prefix = str(tmp_path) + "/"
    m.write_pos(ph, out, prefix, 7, 0.1)
    names = (["rho", "ux"] + (["uy"] if dim == 2 else []) + (["g"] if ph in (MULTIFLUID, STIFFENED) else []) +
             (["pinf"] if ph == STIFFENED else []) + ["p"] + (["phic", "phinc"] if ph == PASSIVE else []))
    for n in names:
        assert (tmp_path / f"{n}0000000007.pos").exists()
    lines = (tmp_path / "rho0000000007.pos").read_text().split("\n")
    assert lines[:13] == ["$MeshFormat", "2.1 0 8", "$EndMeshFormat", "$ElementNodeData", "1", '"Rho"', "1",
                          "0.10000000000000000555", "4", "7", "1", str(t["N_E"]), "0"]
    first = lines[13].split(" ")
    assert first[0] == str(t["elem_id"][0]) and first[1] == str(t["N_s"]) and first[-1] == ""  # trailing space
    assert lines[13 + t["N_E"]] == "$EndElementNodeData"
    back, time = m.read_pos(ph, prefix, 7)
    assert time == 0.1
    assert np.array_equal(back, out)


def test_pos_files_carry_the_file_element_ids(tmp_path):
    m = Mesh(2, 1, 4, 3, bc=(FARFIELD,) * 4, nfields=5)
    m.write_gmsh(tmp_path / "m.msh")
    g = Mesh.from_gmsh(tmp_path / "m.msh", nfields=5)
    assert np.array_equal(g.array("elem_id"), m.array("elem_id"))
    nb = int((m.array("face_tag") != 0).sum())
    assert np.array_equal(m.array("elem_id"), nb + 1 + np.arange(12))


def test_pos_files_that_do_not_fit_are_refused(tmp_path):
    m = Mesh(2, 2, 6, 4, bc=(FARFIELD,) * 4, nfields=5)
    other = Mesh(2, 2, 6, 5, bc=(FARFIELD,) * 4, nfields=5)
    out = oracle.Problem(m.tables(), MULTIFLUID, ROE).format(problems.riemann_state(m, MULTIFLUID, seed=1))
    prefix = str(tmp_path) + "/"
    m.write_pos(MULTIFLUID, out, prefix, 3, 0.5)
    with pytest.raises(DgxError, match="number of elements"):
        other.read_pos(MULTIFLUID, prefix, 3)
    with pytest.raises(DgxError, match="no file"):
        m.read_pos(MULTIFLUID, prefix, 4)
