"""Mixed-precision modes of the tile kernel (FP32-MFMA and split-BF16 map/integrate).

The map to the Gauss points and the integration run in FP32 on MFMA (mode 1)
or as split-BF16 products with FP32 accumulation (mode 2); the fluxes, the
alpha terms, the Riemann solver, Dt*Minv, the RK4 accumulation and the
limiter stay FP64.  Bars against the FP64 oracle, one RHS (relative, max
norm) and 12 HR-limited RK4 steps (relative L2 per field): FP32 2e-6 and
1e-6; split-BF16 5e-5 and 1e-5 (measured 1.9e-5 and 2.0e-6: the dropped
lo*lo term and BF16 rounding of the low parts leave ~2^-16 per product).
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


# precision mode: (RHS bar, trajectory bar, RHS floor showing the reduced precision really ran)
BARS = {1: (2e-6, 1e-6, 1e-10), 2: (5e-5, 1e-5, 1e-7)}


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("ph,fl", [(SINGLEFLUID, ROE), (MULTIFLUID, ROE), (MULTIFLUID, RUSANOV)])
def test_mixed_precision_rhs(device, ph, fl, prec):
    mesh = _mesh(ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(2, 2, ph, fl), mesh)
    solver.set_precision(prec)
    bar, _, floor = BARS[prec]
    errs = []
    for U0 in (problems.smooth_random_state(mesh, ph, seed=1), problems.oblique_riemann_state(mesh, ph, seed=2)):
        f_ref = prob.rhs(U0)
        U = torch.from_numpy(U0).to(device)
        f = torch.zeros_like(U)
        solver.rhs(U, f)
        torch.cuda.synchronize()
        err = np.abs(f.cpu().numpy() - f_ref).max() / np.abs(f_ref).max()
        assert err < bar, err
        errs.append(err)
    assert max(errs) > floor, errs  # the contractions really ran at the reduced precision


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_limited_trajectory(device, prec):
    ph, fl = MULTIFLUID, ROE
    mesh = _mesh(ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(2, 2, ph, fl), mesh)
    solver.set_precision(prec)
    U0 = problems.oblique_riemann_state(mesh, ph, seed=5)
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / 5)
    U_ref, _, _, _ = prob.rk4(U0, 0, dt, nsteps=12, limiter=LIMIT_HR)
    U = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(U, 0, dt, nsteps=12, limiter=LIMIT_HR)
    for l2, li in per_field_rel(U.cpu().numpy(), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < BARS[prec][1], l2


def test_mixed_precision_needs_the_tile_path(device):
    mesh = Mesh(2, 2, 12, 7, 0.0, 1.0, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    for prec in (1, 2):
        with pytest.raises(DgxError):
            solver.set_precision(prec)
