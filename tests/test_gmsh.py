"""Gmsh 2.2 mesh input and output (simpleMesh::load, simpleMesh.cc:141-204;
BuildInterfaces, simpleMesh.cc:770-989).

Pinned by the structured builder, itself pinned against the reference's
conventions in test_setup_tables.py: a mesh written to .msh and read back
gives every table bit for bit, including the face order (sorted node-id
keys), the periodic partners (found by geometry on read) and the boundary
tags; a file with scrambled node ids and element order describes the same
mesh up to the permutation (neighbours exactly, the oracle's right-hand side
to rounding).  No GPU.
"""
import numpy as np
import pytest

import oracle
from dg_amd import FARFIELD, MULTIFLUID, PERIODIC, REFLECTIVE, ROE, DgxError, Mesh, problems

CASES = [
    (1, p, 16, 1, bc) for p in (1, 2, 3) for bc in [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC)]
] + [
    (2, p, 6, 4, bc) for p in (1, 2, 3)
    for bc in [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (FARFIELD, REFLECTIVE, PERIODIC, PERIODIC),
               (PERIODIC,) * 4, (REFLECTIVE,) * 4]
]


def _mesh(d, p, nx, ny, bc):
    if d == 1:
        return Mesh(1, p, nx, 1, -1.0, 1.0, bc=bc, nfields=3 + 1)
    return Mesh(2, p, nx, ny, 0.0, 1.5, -0.25, 0.75, bc=bc, nfields=5)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d_p{c[1]}_{'-'.join(map(str, c[4]))}")
def test_write_then_read_gives_the_same_tables(tmp_path, case):
    d, p, nx, ny, bc = case
    m = _mesh(d, p, nx, ny, bc)
    f = tmp_path / "m.msh"
    m.write_gmsh(f)
    g = Mesh.from_gmsh(f, nfields=m.sizes["N_F"])
    a, b = m.tables(), g.tables()
    for k in a:
        if k in ("nx", "ny"):
            continue  # structured sizes: a file mesh has none
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert not g.tileable  # a file mesh runs the general kernels


def _read_msh(path):
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    nodes = [l.split() for l in lines[i + 2:i + 2 + n]]
    j = lines.index("$Elements")
    m = int(lines[j + 1])
    elems = [l.split() for l in lines[j + 2:j + 2 + m]]
    return nodes, elems


def _write_msh(path, nodes, elems):
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(nodes))
        for n in nodes:
            f.write(" ".join(n) + "\n")
        f.write("$EndNodes\n$Elements\n%d\n" % len(elems))
        for e in elems:
            f.write(" ".join(e) + "\n")
        f.write("$EndElements\n")


@pytest.mark.parametrize("bc", [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (REFLECTIVE, FARFIELD, PERIODIC, PERIODIC)])
def test_scrambled_numbering_is_the_same_mesh(tmp_path, bc):
    m = _mesh(2, 2, 6, 4, bc)
    f = tmp_path / "m.msh"
    m.write_gmsh(f)
    nodes, elems = _read_msh(f)
    rng = np.random.default_rng(7)
    # new node ids
    new_id = rng.permutation(len(nodes)) + 1
    ren = {int(n[0]): int(new_id[k]) for k, n in enumerate(nodes)}
    nodes2 = [[str(ren[int(n[0])])] + n[1:] for n in nodes]
    nodes2 = [nodes2[k] for k in rng.permutation(len(nodes2))]
    vol = [e for e in elems if e[1] == "10"]
    bnd = [e for e in elems if e[1] != "10"]
    perm = rng.permutation(len(vol))  # scrambled element k = original element perm[k]

    def renum(e):
        nt = int(e[2])
        return e[:3 + nt] + [str(ren[int(x)]) for x in e[3 + nt:]]

    elems2 = [renum(e) for e in (bnd[::-1] + [vol[k] for k in perm])]
    f2 = tmp_path / "s.msh"
    _write_msh(f2, nodes2, elems2)
    g = Mesh.from_gmsh(f2, nfields=5)
    tm, tg = m.tables(), g.tables()
    assert tg["N_E"] == tm["N_E"] and tg["M_T"] == tm["M_T"] and tg["Dx"] == tm["Dx"]
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    nb_m, nb_g = tm["neighbors"].reshape(-1, 4), tg["neighbors"].reshape(-1, 4)
    want = np.where(nb_m[perm] >= 0, inv[np.maximum(nb_m[perm], 0)], nb_m[perm])
    assert np.array_equal(nb_g, want)
    # the oracle's right-hand side, permuted with the elements
    U = problems.oblique_riemann_state(m, MULTIFLUID, seed=3)
    nfs = 5 * tm["N_s"]
    f_m = oracle.Problem(tm, MULTIFLUID, ROE).rhs(U).reshape(-1, nfs)
    f_g = oracle.Problem(tg, MULTIFLUID, ROE).rhs(np.ascontiguousarray(U.reshape(-1, nfs)[perm]).reshape(-1))
    assert np.abs(f_g.reshape(-1, nfs) - f_m[perm]).max() <= 1e-12 * np.abs(f_m).max()


def test_bad_files_are_refused(tmp_path):
    m = _mesh(2, 2, 4, 4, (FARFIELD,) * 4)
    f = tmp_path / "m.msh"
    m.write_gmsh(f)
    nodes, elems = _read_msh(f)
    # an interior node moved: not affine
    vol = [e for e in elems if e[1] == "10"]
    centre = vol[0][-1]
    nodes_bad = [n if n[0] != centre else [n[0], str(float(n[1]) + 0.01), n[2], n[3]] for n in nodes]
    _write_msh(tmp_path / "a.msh", nodes_bad, elems)
    with pytest.raises(DgxError, match="affine"):
        Mesh.from_gmsh(tmp_path / "a.msh", nfields=5)
    # a boundary element missing: a face without a second element
    _write_msh(tmp_path / "b.msh", nodes, [e for e in elems if e[1] == "10"] + [e for e in elems if e[1] != "10"][1:])
    with pytest.raises(DgxError, match="boundary"):
        Mesh.from_gmsh(tmp_path / "b.msh", nfields=5)
    # not Gmsh 2.2 ASCII
    (tmp_path / "c.msh").write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
    with pytest.raises(DgxError, match="2.x ASCII"):
        Mesh.from_gmsh(tmp_path / "c.msh", nfields=5)
