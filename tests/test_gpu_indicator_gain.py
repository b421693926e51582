"""What the trained indicator gains over the method without it.

The network is trained on the GPU (dg_amd.indicator: seeded Riemann mosaics,
random interface problems and random shocked bubbles labelled by the
reference sensor), then judged by the plain sensor on a held-out two-fluid
multimode interface hit by a shock and on the shock-bubble (Mach 1.22,
r = 0.25, rho = 0.138, gamma = 1.667) on a mesh the training did not use,
limited by the network itself: no missed shock cell, few missed flags, and
only a small fraction of the cells flagged (the global HR limiter limits
every cell).  On a smooth two-fluid flow the learned limiter leaves the
unlimited solution (almost) untouched, while the global HR limiter clips it.
"""
import numpy as np
import pytest
import torch

from dg_amd import (LIMIT_HR, LIMIT_LEARNED, LIMIT_NONE, MULTIFLUID, PERIODIC, ROE, Config, Mesh, Solver, indicator,
                    problems)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(device):
    data = indicator.merge(indicator.generate(n_sub=32, sub=8), indicator.interface_data(8),
                           indicator.bubble_data(8), indicator.smooth_data(8))
    net, losses, thr = indicator.train(data, epochs=30)
    assert losses[-1] < 0.2 * losses[0], losses
    return net, thr


def test_held_out_interface_judged_by_the_sensor(device, trained):
    net, thr = trained
    mesh = indicator.multimode_interface_mesh(256, 128)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
    solver.set_indicator(net, thr)
    U = torch.from_numpy(indicator.multimode_interface_state(mesh, seed=4242)).to(device)
    NE = mesh.sizes["N_E"]
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    flags = torch.zeros_like(sens)
    for steps in (0, 10):
        if steps:
            solver.rk4_integrate(U, 1, 0.3 * mesh.dx / 5, nsteps=steps, tf=1e30, limiter=LIMIT_HR)
        solver.sensing(U, sens)
        solver.indicator_flags(U, flags)
        torch.cuda.synchronize()
        c = indicator.confusion(flags.cpu().numpy(), sens.cpu().numpy())
        assert c["missed_shock"] == 0, c
        assert c["missed"] <= 0.05 * c["sensor_flagged"], c
        assert c["indicator_flagged"] <= 0.1 * c["cells"], c


def test_shock_bubble_limited_by_the_network(device, trained):
    net, thr = trained
    mesh = problems.shock_bubble_mesh(384, 192, 2)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
    solver.set_indicator(net, thr)
    U = torch.from_numpy(problems.shock_bubble_state(mesh)).to(device)
    NE = mesh.sizes["N_E"]
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    flags = torch.zeros_like(sens)
    done = 0
    for steps in (10, 40, 80):
        solver.rk4_integrate(U, 1, 0.3 * mesh.dx / 5, nsteps=steps - done, tf=1e30, limiter=LIMIT_LEARNED)
        done = steps
        solver.sensing(U, sens)
        solver.indicator_flags(U, flags)
        torch.cuda.synchronize()
        c = indicator.confusion(flags.cpu().numpy(), sens.cpu().numpy())
        assert c["shock_cells"] > 0, c
        assert c["missed_shock"] == 0, (steps, c)
        assert c["missed"] <= 0.05 * c["sensor_flagged"], (steps, c)
        assert c["indicator_flagged"] <= 0.1 * c["cells"], (steps, c)


def test_smooth_flow_is_not_limited(device, trained):
    net, thr = trained
    # smooth enough that the sensor flags no cell (checked against the oracle:
    # 0 of 2048 flagged at the thresholds of the training data)
    mesh = Mesh(2, 2, 64, 32, 0.0, 2.0, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)
    U0 = problems.smooth_random_state(mesh, MULTIFLUID, seed=12, amp=0.02)
    out = {}
    for lim in (LIMIT_NONE, LIMIT_HR, LIMIT_LEARNED):
        solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
        solver.set_indicator(net, thr)
        U = torch.from_numpy(U0).to(device)
        solver.rk4_integrate(U, 0, 2e-3, nsteps=10, limiter=lim)
        out[lim] = U.cpu().numpy()
    d_hr = np.linalg.norm(out[LIMIT_HR] - out[LIMIT_NONE])
    d_learned = np.linalg.norm(out[LIMIT_LEARNED] - out[LIMIT_NONE])
    assert d_hr > 1e-6 * np.linalg.norm(out[LIMIT_NONE])
    assert d_learned <= 0.1 * d_hr, (d_learned, d_hr)
