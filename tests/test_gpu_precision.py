"""Mixed-precision mode of the tile kernel (FP32-MFMA map/integrate).

The map to the Gauss points and the integration run in FP32 on MFMA; the
fluxes, the alpha terms, the Riemann solver, Dt*Minv, the RK4 accumulation
and the limiter stay FP64.  Bars: one RHS within 2e-6 relative (max norm) of the FP64
oracle; 12 limited RK4 steps within 1e-6 relative L2 per field.
"""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import (FARFIELD, LIMIT_HR, MULTIFLUID, PERIODIC, REFLECTIVE, ROE, RUSANOV, SINGLEFLUID, Config, DgxError,
                    Mesh, Solver, problems)
from tests.helpers import per_field_rel

pytestmark = pytest.mark.gpu


def _mesh(ph, bc=(FARFIELD, REFLECTIVE, PERIODIC, PERIODIC)):
    return Mesh(2, 2, 32, 16, 0.0, 1.6, 0.0, 0.8, bc=bc, nfields=4 + ph)


@pytest.mark.parametrize("ph,fl", [(SINGLEFLUID, ROE), (MULTIFLUID, ROE), (MULTIFLUID, RUSANOV)])
def test_mixed_precision_rhs(device, ph, fl):
    mesh = _mesh(ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(2, 2, ph, fl), mesh)
    solver.set_precision(1)
    for U0 in (problems.smooth_random_state(mesh, ph, seed=1), problems.oblique_riemann_state(mesh, ph, seed=2)):
        f_ref = prob.rhs(U0)
        U = torch.from_numpy(U0).to(device)
        f = torch.zeros_like(U)
        solver.rhs(U, f)
        torch.cuda.synchronize()
        err = np.abs(f.cpu().numpy() - f_ref).max() / np.abs(f_ref).max()
        assert err < 2e-6, err
        assert err > 1e-10, err  # the contractions really ran in FP32


def test_mixed_precision_limited_trajectory(device):
    ph, fl = MULTIFLUID, ROE
    mesh = _mesh(ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(2, 2, ph, fl), mesh)
    solver.set_precision(1)
    U0 = problems.oblique_riemann_state(mesh, ph, seed=5)
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / 5)
    U_ref, _, _, _ = prob.rk4(U0, 0, dt, nsteps=12, limiter=LIMIT_HR)
    U = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(U, 0, dt, nsteps=12, limiter=LIMIT_HR)
    for l2, li in per_field_rel(U.cpu().numpy(), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-6, l2


def test_mixed_precision_needs_the_tile_path(device):
    mesh = Mesh(2, 2, 12, 7, 0.0, 1.0, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    with pytest.raises(DgxError):
        solver.set_precision(1)
