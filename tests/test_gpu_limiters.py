"""The m2, primitive-variable HR and p=0 limiters (m2l, m2i, prl, pri, p0i) on the GPU: bit for
bit the oracle's (which equals the reference's launchers, see
test_oracle_vs_reference.py) for 1D and 2D, p = 1..3, walls and periodic
sides; limited RK4 trajectories within 1e-11 of the oracle's; single-fluid
states are refused (the reference implements both for two-fluid GAMNCON)."""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import (FARFIELD, LIMIT_M2, LIMIT_M2I, LIMIT_P0I, LIMIT_PR, LIMIT_PRI, MULTIFLUID, PERIODIC, REFLECTIVE, ROE,
                    SINGLEFLUID, Config, DgxError, Mesh, Solver, problems)
from tests.helpers import per_field_rel, small_mesh

pytestmark = pytest.mark.gpu

BCS = {1: [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC)],
       2: [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE)]}


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_m2_and_pr_bitwise(device, dim, order):
    for k, bc in enumerate(BCS[dim]):
        mesh = small_mesh(dim, order, MULTIFLUID, bc=bc)
        prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
        solver = Solver(Config(dim, order, MULTIFLUID, ROE), mesh)
        U0 = problems.riemann_state(mesh, MULTIFLUID, seed=40 + k)
        x = np.arange(U0.size, dtype=np.float64)
        U0 = U0 * (1.0 + 0.01 * np.sin(0.37 * x + k))
        th = (0.01, 0.01, 0.01)
        solver.set_sensor_thresholds(th)
        for lim, ref in ((LIMIT_M2, prob.limit_m2), (LIMIT_PR, prob.limit_pr),
                         (LIMIT_PRI, lambda u: prob.limit_pri(u, th)[0]),
                         (LIMIT_M2I, lambda u: prob.limit_m2i(u, th)[0]),
                         (LIMIT_P0I, lambda u: prob.limit_p0i(u, th)[0])):
            U = torch.from_numpy(U0).to(device)
            solver.limit(lim, U)
            torch.cuda.synchronize()
            assert np.array_equal(U.cpu().numpy(), ref(U0)), (lim, bc)


@pytest.mark.parametrize("dim,n", [(1, 256), (2, (32, 16))])
@pytest.mark.parametrize("lim", [LIMIT_M2, LIMIT_M2I, LIMIT_PR, LIMIT_PRI, LIMIT_P0I])
def test_limited_trajectories(device, dim, n, lim):
    if dim == 1:
        mesh = Mesh(1, 2, n, 1, -1.0, 1.0, bc=(FARFIELD, REFLECTIVE), nfields=4)
        U0 = problems.riemann_state(mesh, MULTIFLUID, seed=3)
    else:
        mesh = Mesh(2, 2, n[0], n[1], 0.0, 1.6, 0.0, 0.8, bc=(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=5)
        U0 = problems.oblique_riemann_state(mesh, MULTIFLUID, seed=3)
    prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
    solver = Solver(Config(dim, 2, MULTIFLUID, ROE), mesh)
    th = (0.01, 0.01, 0.01)
    solver.set_sensor_thresholds(th)
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / 5)
    U_ref, _, _, _ = prob.rk4(U0, 0, dt, nsteps=8, limiter=lim, thresholds=th)
    U = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(U, 0, dt, nsteps=8, limiter=lim)
    torch.cuda.synchronize()
    for l2, li in per_field_rel(U.cpu().numpy(), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


def test_single_fluid_is_refused(device):
    mesh = small_mesh(2, 2, SINGLEFLUID)
    solver = Solver(Config(2, 2, SINGLEFLUID, ROE), mesh)
    U = torch.from_numpy(problems.riemann_state(mesh, SINGLEFLUID, seed=1)).to(device)
    for lim in (LIMIT_M2, LIMIT_M2I, LIMIT_PR, LIMIT_PRI, LIMIT_P0I):
        with pytest.raises(DgxError):
            solver.limit(lim, U)


def test_m2i_many_flagged_bitwise(device):
    """More flagged cells than one pass of the m2i grid covers (8192): the
    block loop over the list, bit for bit the oracle."""
    mesh = Mesh(2, 2, 128, 96, 0.0, 1.6, 0.0, 1.2, bc=(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=5)
    prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    U0 = problems.oblique_riemann_state(mesh, MULTIFLUID, seed=5)
    rng = np.random.default_rng(5)
    U0 = U0 * (1.0 + 0.05 * rng.standard_normal(U0.shape))
    th = (0.001, 0.001, 0.001)
    solver.set_sensor_thresholds(th)
    ref, sens = prob.limit_m2i(U0, th)
    assert (sens != 0).sum() > 8192
    U = torch.from_numpy(U0).to(device)
    solver.limit(LIMIT_M2I, U)
    torch.cuda.synchronize()
    assert np.array_equal(U.cpu().numpy(), ref)
