"""CFL runs with the time control on the device (dgx_rk4_integrate mode 1,
limiter 0 or 1): the steps are enqueued in chunks, k_dt_step stops the run
and clamps the last step, and the kernels of steps queued after the stop
return at once.  The result must be the per-step host loop's (rk.cc:140-150)
bit for bit: the same state, step count and final time, also when the final
time falls inside the first chunk (queued steps that must do nothing) and
when a NaN stops the run.
"""
import numpy as np
import pytest
import torch

from dg_amd import LIMIT_HR, MULTIFLUID, RUSANOV, ROE, SINGLEFLUID, Config, DgxError, Solver, problems

pytestmark = pytest.mark.gpu


def _run(solver, U0, device, on, cfl, nsteps=0, tf=0.0, limiter=0):
    solver.set_device_time(on)
    U = torch.from_numpy(np.ascontiguousarray(U0)).to(device)
    steps, T = solver.rk4_integrate(U, 1, cfl, nsteps=nsteps, tf=tf, limiter=limiter)
    torch.cuda.synchronize()
    return U.cpu().numpy(), steps, T


def _same(solver, U0, device, cfl, **kw):
    Ud, nd, Td = _run(solver, U0, device, True, cfl, **kw)
    Uh, nh, Th = _run(solver, U0, device, False, cfl, **kw)
    assert nd == nh and Td == Th, (nd, nh, Td, Th)
    assert np.array_equal(Ud, Uh)
    return nd, Td


@pytest.mark.parametrize("tf,nsteps", [(0.1, 0), (0.1, 9), (1e-3, 0)])
def test_sod_device_time_equals_host_loop(device, tf, nsteps):
    """1D Sod, face + element kernels and the two-launch HR limiter: to the
    final time (several chunks), capped at 9 steps, and a final time reached
    after 2 steps (the rest of the first chunk of 4 is skipped)."""
    mesh = problems.sod_mesh(200, 2, SINGLEFLUID)
    solver = Solver(Config(1, 2, SINGLEFLUID, RUSANOV), mesh)
    cfl = 0.4 * mesh.dx / 5
    n, T = _same(solver, problems.sod_state(mesh, SINGLEFLUID), device, cfl, nsteps=nsteps, tf=tf,
                 limiter=LIMIT_HR)
    if nsteps:
        assert n == nsteps
    else:
        assert abs(T - tf) <= 1e-14
    if tf == 1e-3:
        assert n == 2


@pytest.mark.parametrize("limiter", [0, LIMIT_HR])
def test_shock_bubble_device_time_equals_host_loop(device, limiter):
    """2D two-fluid shock-bubble on the tile kernel (with the one-launch HR
    limiter for limiter 1), to a final time about a dozen CFL steps away."""
    mesh = problems.shock_bubble_mesh(64, 32, 2)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    cfl = 0.4 * mesh.dx / 5
    U0 = problems.shock_bubble_state(mesh)
    n, T = _same(solver, U0, device, cfl, tf=0.009, limiter=limiter)
    assert abs(T - 0.009) <= 1e-14 and n > 4


def test_nan_stops_both_loops(device):
    """A NaN in the state: Dt is NaN at the first step and both schedules
    stop with DGX_EDT (rk.cc:147-148), leaving the same state."""
    mesh = problems.sod_mesh(64, 2, SINGLEFLUID)
    solver = Solver(Config(1, 2, SINGLEFLUID, RUSANOV), mesh)
    U0 = problems.sod_state(mesh, SINGLEFLUID).copy()
    U0.reshape(-1)[7] = np.nan
    cfl = 0.4 * mesh.dx / 5
    out = []
    for on in (True, False):
        solver.set_device_time(on)
        U = torch.from_numpy(U0.copy()).to(device)
        with pytest.raises(DgxError, match=r"\(-4\)"):
            solver.rk4_integrate(U, 1, cfl, tf=0.1, limiter=LIMIT_HR)
        torch.cuda.synchronize()
        out.append(U.cpu().numpy())
    assert np.array_equal(out[0], out[1], equal_nan=True)
