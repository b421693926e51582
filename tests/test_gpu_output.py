"""Solution output and restart on the GPU: the formatter kernel equals the
oracle's (itself the reference's Lformater bit for bit) both ways, a
checkpoint read back through the inverse formatter gives the oracle's
restart state bit for bit, and a run restarted from a checkpoint follows the
uninterrupted run to rounding (the restart formatter's own)."""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import (FARFIELD, LIMIT_HR, MULTIFLUID, PASSIVE, PERIODIC, REFLECTIVE, ROE, SINGLEFLUID, STIFFENED, Config,
                    Mesh, Solver, nfields, problems)
from tests.helpers import per_field_rel

pytestmark = pytest.mark.gpu


def _setup(dim, ph):
    nf = nfields(dim, ph)
    if dim == 1:
        m = Mesh(1, 2, 64, 1, -1.0, 1.0, bc=(FARFIELD, REFLECTIVE), nfields=nf)
    else:
        m = Mesh(2, 2, 16, 8, 0.0, 1.6, 0.0, 0.8, bc=(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=nf)
    return m, oracle.Problem(m.tables(), ph, ROE), Solver(Config(dim, 2, ph, ROE), m)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("ph", [SINGLEFLUID, MULTIFLUID, PASSIVE, STIFFENED])
def test_formatter_and_checkpoint(device, tmp_path, dim, ph):
    m, prob, solver = _setup(dim, ph)
    U0 = problems.riemann_state(m, ph, seed=8)
    U = torch.from_numpy(U0).to(device)
    out = torch.zeros_like(U)
    solver.format(U, out)
    torch.cuda.synchronize()
    ref = prob.format(U0)
    assert np.array_equal(out.cpu().numpy(), ref)
    back = torch.zeros_like(U)
    solver.format(back, out, inverse=True)
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy(), prob.unformat(ref))
    prefix = str(tmp_path) + "/"
    solver.checkpoint(U, prefix, 12, 0.25)
    U2 = torch.zeros_like(U)
    assert solver.restart(prefix, 12, U2) == 0.25
    assert np.array_equal(U2.cpu().numpy(), prob.unformat(ref))


def test_restarted_run_follows_the_uninterrupted_one(device, tmp_path):
    m, prob, solver = _setup(2, MULTIFLUID)
    U0 = problems.oblique_riemann_state(m, MULTIFLUID, seed=9)
    dt = 0.3 * prob.dt_from_cfl(U0, m.dx / 5)
    Ua = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(Ua, 0, dt, nsteps=8, limiter=LIMIT_HR)
    Ub = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(Ub, 0, dt, nsteps=4, limiter=LIMIT_HR)
    prefix = str(tmp_path) + "/"
    solver.checkpoint(Ub, prefix, 4, 4 * dt)
    Uc = torch.zeros_like(Ub)
    assert solver.restart(prefix, 4, Uc) == 4 * dt
    solver.rk4_integrate(Uc, 0, dt, nsteps=4, limiter=LIMIT_HR)
    torch.cuda.synchronize()
    for l2, li in per_field_rel(Uc.cpu().numpy(), Ua.cpu().numpy(), prob.nf, prob.cfg.N_s):
        assert l2 < 1e-12 and li < 1e-12, (l2, li)
