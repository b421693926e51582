"""GPU checks of the learned indicator against its CPU restatement.

Features (float32 of FP64 arithmetic) within 1e-6; logits of the BF16 MFMA
forward within 0.05 absolute of the oracle's BF16 forward (a BF16 rounding
step of an activation moves z by ~1e-2 at most; a wrong operand layout moves
it by O(1)); gradients within 2% relative L2 per parameter block; training
is deterministic and learns a separable rule; the learned limiter applies
the HR sweeps exactly (bit for bit) on the flags the network produced.
"""
import numpy as np
import pytest
import torch

import oracle
import oracle.mlp as om
from dg_amd import LIMIT_LEARNED, MULTIFLUID, PERIODIC, REFLECTIVE, FARFIELD, ROE, Config, Indicator, Mesh, Solver, problems

pytestmark = pytest.mark.gpu


def _solver(nx=16, ny=8, bc=(FARFIELD, REFLECTIVE, PERIODIC, PERIODIC)):
    mesh = Mesh(2, 2, nx, ny, 0.0, 1.6, 0.0, 0.8, bc=bc, nfields=5)
    return mesh, Solver(Config(2, 2, MULTIFLUID, ROE), mesh)


def test_features_match_oracle(device):
    mesh, solver = _solver()
    t = mesh.tables()
    net = Indicator(seed=1)
    solver.set_indicator(net, 0.0)
    vinv = om.legendre_vinv(t["ref_nodes"])
    for U0 in (problems.oblique_riemann_state(mesh, MULTIFLUID, seed=3),
               problems.smooth_random_state(mesh, MULTIFLUID, seed=4)):
        U = torch.from_numpy(U0).to(device)
        X = torch.zeros((t["N_E"], 64), dtype=torch.float32, device=device)
        solver.features(U, X)
        torch.cuda.synchronize()
        ref = om.features(U0, t["neighbors"], vinv)
        assert np.allclose(X.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)


def test_fused_indicator_equals_features_then_forward(device):
    mesh, solver = _solver()
    t = mesh.tables()
    net = Indicator(seed=4)
    solver.set_indicator(net, 0.25)
    U = torch.from_numpy(problems.oblique_riemann_state(mesh, MULTIFLUID, seed=6)).to(device)
    X = torch.zeros((t["N_E"], 64), dtype=torch.float32, device=device)
    z1 = torch.zeros(t["N_E"], dtype=torch.float32, device=device)
    z2 = torch.zeros_like(z1)
    flags = torch.zeros(t["N_E"], dtype=torch.int32, device=device)
    solver.features(U, X)
    net.forward(X, z1)
    solver.indicator_flags(U, flags, z2)
    torch.cuda.synchronize()
    assert np.array_equal(z1.cpu().numpy(), z2.cpu().numpy())
    wall = (t["neighbors"].reshape(-1, 4) < 0).any(axis=1)
    assert np.array_equal(flags.cpu().numpy(), ((z2.cpu().numpy() > 0.25) | wall).astype(np.int32))


def _random_inputs(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 64), dtype=np.float32)
    X[:, :49] = rng.standard_normal((n, 49)).astype(np.float32)
    return X


def test_forward_matches_oracle(device):
    net = Indicator(seed=7)
    P = net.params
    X = _random_inputs(1000, 2)
    z = torch.zeros(1000, dtype=torch.float32, device=device)
    net.forward(torch.from_numpy(X).to(device), z)
    torch.cuda.synchronize()
    zr, _ = om.forward(P, X)
    d = np.abs(z.cpu().numpy() - zr)
    assert d.max() < 0.05, d.max()
    assert np.mean(d > 1e-3) < 0.05


def test_gradient_matches_oracle(device):
    net = Indicator(seed=11)
    P = net.params
    X = _random_inputs(256, 5)
    rng = np.random.default_rng(6)
    y = (rng.random(256) < 0.3).astype(np.float32)
    g = torch.zeros(net.n_params + 1, dtype=torch.float32, device=device)
    net.gradient(torch.from_numpy(X).to(device), torch.from_numpy(y).to(device), g, pos_weight=2.5)
    torch.cuda.synchronize()
    gg = g.cpu().numpy().astype(np.float64)
    gr = om.gradient(P, X, y.astype(np.float64), 2.5)
    assert abs(gg[-1] - om.loss(P, X, y, 2.5)) <= 1e-3 * abs(om.loss(P, X, y, 2.5))
    for a, b in ((om.W1, om.B1), (om.B1, om.W2), (om.W2, om.B2), (om.B2, om.W3), (om.W3, om.B3), (om.B3, om.N_PARAM)):
        err = np.linalg.norm(gg[a:b] - gr[a:b]) / max(np.linalg.norm(gr[a:b]), 1e-12)
        assert err < 0.02, (a, b, err)


def test_training_is_deterministic_and_learns(device):
    rng = np.random.default_rng(9)
    X = _random_inputs(4096, 9)
    y = ((X[:, 45] + 0.5 * X[:, 3]) > 0.4).astype(np.float32)
    Xd, yd = torch.from_numpy(X).to(device), torch.from_numpy(y).to(device)
    runs = []
    for _ in range(2):
        net = Indicator(seed=3)
        losses = net.train(Xd, yd, epochs=12, batch=256, lr=3e-3, pos_weight=1.0, seed=5)
        runs.append((net.params, losses))
        z = torch.zeros(4096, dtype=torch.float32, device=device)
        net.forward(Xd, z)
        torch.cuda.synchronize()
        acc = np.mean((z.cpu().numpy() > 0) == (y > 0.5))
        assert acc > 0.95, acc
    assert np.array_equal(runs[0][0], runs[1][0])
    losses = runs[0][1]
    assert losses[-1] < 0.5 * losses[0], losses


def test_learned_limiter_is_exact_on_its_flags(device):
    mesh, solver = _solver()
    t = mesh.tables()
    prob = oracle.Problem(t, MULTIFLUID, ROE)
    net = Indicator(seed=2)
    U0 = problems.oblique_riemann_state(mesh, MULTIFLUID, seed=8)
    Ud = torch.from_numpy(U0).to(device)
    z = torch.zeros(t["N_E"], dtype=torch.float32, device=device)
    flags = torch.zeros(t["N_E"], dtype=torch.int32, device=device)
    solver.set_indicator(net, 0.0)
    solver.indicator_flags(Ud, flags, z)
    torch.cuda.synchronize()
    zz = z.cpu().numpy()
    thr = float(np.median(zz))  # flag about half of the cells
    solver.set_indicator(net, thr)
    solver.indicator_flags(Ud, flags)
    solver.limit(LIMIT_LEARNED, Ud)
    torch.cuda.synchronize()
    f = flags.cpu().numpy()
    assert 0 < f.sum() < t["N_E"]
    wall = (t["neighbors"].reshape(-1, 4) < 0).any(axis=1)  # flagged by rule, as by the sensor
    assert np.array_equal(f, ((zz > thr) | wall).astype(np.int32))
    assert np.array_equal(Ud.cpu().numpy(), prob.limit_flagged(U0, f))


@pytest.mark.parametrize("prec", [0, 1])
def test_stage_emitted_means_give_the_same_run(device, prec):
    """The tile kernel's epilogue writes the cell means the indicator reads
    (dgx_solver_set_fused_means, on by default) and the sensor's cell
    averages the shock guard reads: 6 RK4 steps limited by the learned
    indicator on the shock-bubble at 128x64 give the same state bit for bit
    with and without it, in FP64 and with the FP32 contractions.  An
    untrained seeded network with the threshold at the median logit of the
    initial state flags a large share of the cells."""
    from dg_amd import PATH_TILED

    mesh = problems.shock_bubble_mesh(128, 64)
    U0 = problems.shock_bubble_state(mesh, xs=0.24)
    net = Indicator(seed=3)
    out = []
    from dg_amd import indicator

    for fused in (True, False):
        solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
        assert solver.path == PATH_TILED
        solver.set_precision(prec)
        # the sensor thresholds switch the shock guard on: the stages also
        # write the sensor's cell averages it reads
        solver.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
        solver.set_indicator(net, 0.0)
        Ud = torch.from_numpy(U0).to(device)
        z = torch.zeros(mesh.sizes["N_E"], dtype=torch.float32, device=device)
        flags = torch.zeros(mesh.sizes["N_E"], dtype=torch.int32, device=device)
        solver.indicator_flags(Ud, flags, z)
        torch.cuda.synchronize()
        thr = float(np.median(z.cpu().numpy()))
        solver.set_indicator(net, thr)
        solver.set_fused_means(fused)
        dt_dev = torch.zeros(1, dtype=torch.float64, device=device)
        solver.dt_from_cfl(Ud, 0.4 * mesh.dx / 5, dt_dev)
        dt = float(dt_dev.item())
        solver.rk4_integrate(Ud, 0, dt, nsteps=6, limiter=LIMIT_LEARNED)
        solver.indicator_flags(Ud, flags)
        torch.cuda.synchronize()
        out.append((Ud.cpu().numpy(), flags.cpu().numpy()))
    assert 0 < out[0][1].sum() < out[0][1].size
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    assert not np.array_equal(out[0][0], U0)


def test_shock_guard_holds_every_sensor_shock_cell(device):
    """With a threshold no logit reaches, the indicator's flags are the
    boundary cells and the shock guard alone: on the shock-bubble at 256x128
    (shock, reflected and transmitted shocks after 40 HR-limited steps) they
    contain every cell the sensor marks as a shock, bit-identical decisions
    from the same cell averages; without the guard the untrained network
    flags almost none of them."""
    from dg_amd import LIMIT_HR, indicator

    mesh = problems.shock_bubble_mesh(256, 128)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
    solver.set_indicator(Indicator(seed=5), 1e30)
    U = torch.from_numpy(problems.shock_bubble_state(mesh, xs=0.24)).to(device)
    solver.rk4_integrate(U, 1, 0.4 * mesh.dx / 5, nsteps=40, tf=1e30, limiter=LIMIT_HR)
    NE = mesh.sizes["N_E"]
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    flags = torch.zeros_like(sens)
    solver.sensing(U, sens)
    solver.indicator_flags(U, flags)
    torch.cuda.synchronize()
    c = indicator.confusion(flags.cpu().numpy(), sens.cpu().numpy())
    print("guard alone:", c)
    assert c["shock_cells"] > 100 and c["missed_shock"] == 0, c
    solver.set_shock_guard(False)
    solver.indicator_flags(U, flags)
    torch.cuda.synchronize()
    c0 = indicator.confusion(flags.cpu().numpy(), sens.cpu().numpy())
    print("without the guard:", c0)
    assert c0["missed_shock"] > 0.9 * c0["shock_cells"], c0
