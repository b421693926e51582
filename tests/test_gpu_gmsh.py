"""The engine on a mesh read from a Gmsh file with scrambled node ids and
element order (face + element kernels): RK4 steps with the HR limiter follow
the oracle on the same tables to 1e-11 per field."""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import LIMIT_HR, MULTIFLUID, PATH_GENERAL, PERIODIC, REFLECTIVE, FARFIELD, ROE, Config, Mesh, Solver, problems
from tests.helpers import per_field_rel
from tests.test_gmsh import _mesh, _read_msh, _write_msh

pytestmark = pytest.mark.gpu


def test_rk4_on_a_scrambled_gmsh_mesh(device, tmp_path):
    m = _mesh(2, 2, 8, 6, (PERIODIC, PERIODIC, REFLECTIVE, FARFIELD))
    f = tmp_path / "m.msh"
    m.write_gmsh(f)
    nodes, elems = _read_msh(f)
    rng = np.random.default_rng(11)
    new_id = rng.permutation(len(nodes)) + 1
    ren = {int(n[0]): int(new_id[k]) for k, n in enumerate(nodes)}
    nodes2 = [[str(ren[int(n[0])])] + n[1:] for n in nodes]
    vol = [e for e in elems if e[1] == "10"]
    bnd = [e for e in elems if e[1] != "10"]
    perm = rng.permutation(len(vol))
    renum = lambda e: e[:3 + int(e[2])] + [str(ren[int(x)]) for x in e[3 + int(e[2]):]]
    _write_msh(tmp_path / "s.msh", nodes2, [renum(e) for e in bnd + [vol[k] for k in perm]])
    g = Mesh.from_gmsh(tmp_path / "s.msh", nfields=5)
    t = g.tables()
    prob = oracle.Problem(t, MULTIFLUID, ROE)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), g)
    assert solver.path == PATH_GENERAL
    nfs = 5 * t["N_s"]
    U0 = problems.oblique_riemann_state(m, MULTIFLUID, seed=4).reshape(-1, nfs)[perm].reshape(-1)
    U0 = np.ascontiguousarray(U0)
    dt = 0.3 * prob.dt_from_cfl(U0, g.dx / 5)
    U_ref, _, _, _ = prob.rk4(U0, 0, dt, nsteps=6, limiter=LIMIT_HR)
    U = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(U, 0, dt, nsteps=6, limiter=LIMIT_HR)
    torch.cuda.synchronize()
    for l2, li in per_field_rel(U.cpu().numpy(), U_ref, prob.nf, t["N_s"]):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)
