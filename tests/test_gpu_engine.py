"""GPU parity of the fused HIP engine against the CPU oracle.

Tolerances: the FP64 right-hand side differs from the oracle only by FMA
contraction and summation association (relative 1e-12 on one RHS); N-step
trajectories must stay within 1e-11 relative L2/Linf per conserved field;
the limiter and sensor kernels are compiled without contraction and must be
bit-identical to the oracle on identical inputs.
"""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import (FARFIELD, LIMIT_HR, LIMIT_HRI, MULTIFLUID, PERIODIC, REFLECTIVE, ROE, RUSANOV, SINGLEFLUID,
                    Config, Mesh, Solver, problems)
from tests.helpers import CONFIGS, CONFIGS_PS, cfg_id, per_field_rel, rel_err, small_mesh

pytestmark = pytest.mark.gpu


def dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


BCS_1D = [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC), (REFLECTIVE, FARFIELD)]
BCS_2D = [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (FARFIELD, REFLECTIVE, PERIODIC, PERIODIC),
          (REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE)]


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_rhs_matches_oracle(device, c, order):
    d, ph, fl = c
    for k, bc in enumerate(BCS_1D if d == 1 else BCS_2D):
        mesh = small_mesh(d, order, ph, bc=bc)
        prob = oracle.Problem(mesh.tables(), ph, fl)
        solver = Solver(Config(d, order, ph, fl), mesh)
        for U in (problems.smooth_random_state(mesh, ph, seed=k), problems.riemann_state(mesh, ph, seed=k + 7)):
            f_ref = prob.rhs(U)
            Ud = dev(U, device)
            fd = torch.zeros_like(Ud)
            solver.rhs(Ud, fd)
            f = host(fd)
            scale = np.abs(f_ref).max()
            assert np.abs(f - f_ref).max() <= 1e-12 * scale, (bc, np.abs(f - f_ref).max() / scale)


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
def test_rk4_fixed_dt_trajectory(device, c):
    d, ph, fl = c
    order = 2
    mesh = small_mesh(d, order, ph, n=(24, 12))
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(d, order, ph, fl), mesh)
    U0 = problems.smooth_random_state(mesh, ph, seed=3)
    dt = 0.2 * prob.dt_from_cfl(U0, mesh.dx / (2 * order + 1))
    nsteps = 25
    U_ref, n, T, _ = prob.rk4(U0, 0, dt, nsteps=nsteps)
    Ud = dev(U0, device)
    steps, Tg = solver.rk4_integrate(Ud, 0, dt, nsteps=nsteps)
    U = host(Ud)
    assert steps == nsteps
    for l2, li in per_field_rel(U, U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


@pytest.mark.parametrize("c", [x for x in CONFIGS if x[0] == 1], ids=cfg_id)
def test_cfl_mode_sod_with_hr_limiter(device, c):
    """Sod tube to a final time with CFL steps and the global HR limiter."""
    d, ph, fl = c
    order = 2
    mesh = problems.sod_mesh(nx=80, order=order, physics=ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(1, order, ph, fl), mesh)
    U0 = problems.sod_state(mesh, ph)
    cfl = 0.4 * mesh.dx / (2 * order + 1)
    tf = 0.05
    U_ref, n_ref, T_ref, dts = prob.rk4(U0, 1, cfl, tf=tf, limiter=LIMIT_HR)
    Ud = dev(U0, device)
    steps, T = solver.rk4_integrate(Ud, 1, cfl, tf=tf, limiter=LIMIT_HR)
    U = host(Ud)
    assert steps == n_ref and abs(T - T_ref) <= 1e-14
    for l2, li in per_field_rel(U, U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


@pytest.mark.parametrize("c", [x for x in CONFIGS if x[0] == 2], ids=cfg_id)
def test_limited_2d_trajectory_general_path(device, c):
    """Oblique 2D Riemann problem with the HR limiter on a mesh the tile
    kernel does not take (12 x 7 elements): face + element kernels."""
    d, ph, fl = c
    order = 2
    mesh = small_mesh(d, order, ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(d, order, ph, fl), mesh)
    U0 = problems.oblique_riemann_state(mesh, ph, seed=3)
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / (2 * order + 1))
    U_ref, n, T, _ = prob.rk4(U0, 0, dt, nsteps=12, limiter=LIMIT_HR)
    Ud = dev(U0, device)
    solver.rk4_integrate(Ud, 0, dt, nsteps=12, limiter=LIMIT_HR)
    for l2, li in per_field_rel(host(Ud), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


def _stepped_riemann(mesh, prob, ph, seed):
    """A Riemann state with a smooth 1% multiplicative perturbation in every
    field (limiter and sensor input: discontinuous, non-polynomial data)."""
    U0 = problems.riemann_state(mesh, ph, seed=seed)
    x = np.arange(U0.size, dtype=np.float64)
    return U0 * (1.0 + 0.01 * np.sin(0.37 * x + seed))


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_hr_limiter_bitwise(device, dim, order):
    ph, fl = MULTIFLUID, ROE
    for k, bc in enumerate(BCS_1D if dim == 1 else BCS_2D):
        mesh = small_mesh(dim, order, ph, bc=bc)
        prob = oracle.Problem(mesh.tables(), ph, fl)
        solver = Solver(Config(dim, order, ph, fl), mesh)
        U = _stepped_riemann(mesh, prob, ph, seed=20 + k)
        U_ref = prob.limit_hr(U)
        Ud = dev(U, device)
        solver.limit(LIMIT_HR, Ud)
        assert np.array_equal(host(Ud), U_ref)
        assert not np.array_equal(U_ref, U)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_sensors_and_hri_bitwise(device, dim, order):
    ph, fl = MULTIFLUID, ROE
    for k, bc in enumerate(BCS_1D if dim == 1 else BCS_2D):
        mesh = small_mesh(dim, order, ph, bc=bc)
        prob = oracle.Problem(mesh.tables(), ph, fl)
        solver = Solver(Config(dim, order, ph, fl), mesh)
        U = _stepped_riemann(mesh, prob, ph, seed=40 + k)
        for th in [(0.01,), (0.01, 0.01), (0.02, 0.005, 0.01), (1e-4, 1e-4, 1e-4)]:
            sens_ref, uavg_ref = prob.sensing(U, th)
            solver.set_sensor_thresholds(th)
            Ud = dev(U, device)
            sd = torch.zeros(prob.cfg.N_E, dtype=torch.int32, device=device)
            ua = torch.zeros(prob.cfg.N_E * prob.nf, dtype=torch.float64, device=device)
            solver.sensing(Ud, sd, ua)
            assert np.array_equal(host(ua), uavg_ref)
            assert np.array_equal(host(sd), sens_ref), th
            U_ref, _ = prob.limit_hri(U, th)
            solver.limit(LIMIT_HRI, Ud)
            assert np.array_equal(host(Ud), U_ref)


def test_dt_from_cfl(device):
    for c in CONFIGS + CONFIGS_PS:
        d, ph, fl = c
        mesh = small_mesh(d, 2, ph)
        prob = oracle.Problem(mesh.tables(), ph, fl)
        solver = Solver(Config(d, 2, ph, fl), mesh)
        U = problems.riemann_state(mesh, ph, seed=5)
        dt = torch.zeros(1, dtype=torch.float64, device=device)
        solver.dt_from_cfl(dev(U, device), 0.37, dt)
        assert host(dt)[0] == prob.dt_from_cfl(U, 0.37)


def test_vortex_full_size_invariants(device):
    """512^2 isentropic vortex (the benchmark workload): free-stream
    preservation of the RHS, conservation of every field over RK4 steps on a
    periodic mesh, and the error against the exact translated vortex."""
    mesh = problems.vortex_mesh(512, 2)
    solver = Solver(Config(2, 2, SINGLEFLUID, ROE), mesh)
    s = mesh.sizes
    # free stream: RHS = 0 to round-off
    Ufs = np.zeros((s["N_E"], 4, s["N_s"]))
    Ufs[:, 0], Ufs[:, 1], Ufs[:, 2], Ufs[:, 3] = 1.0, 1.0, 1.0, 1.0 / 0.4 + 1.0
    Ud = dev(Ufs.ravel(), device)
    fd = torch.zeros_like(Ud)
    solver.rhs(Ud, fd)
    assert host(fd).__abs__().max() < 1e-12
    # conservation: sum_e J_e * sum_g w_g * U(x_g) is constant
    U0 = problems.vortex_state(mesh)
    w = mesh.array("weight")
    phi = mesh.array("phi").reshape(s["N_s"], s["N_G"]).T  # phi[g, i]
    J = mesh.array("J")

    def totals(U):
        A = U.reshape(s["N_E"], 4, s["N_s"])
        return np.einsum("g,gi,efi,e->f", w, phi, A, J)

    Ud = dev(U0, device)
    dt = 0.3 * mesh.dx / 5 / 2.5
    solver.rk4_integrate(Ud, 0, dt, nsteps=20)
    U1 = host(Ud)
    t0, t1 = totals(U0), totals(U1)
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.abs(t0)), (t0, t1)
    # accuracy against the exact solution at t = 20 dt
    x, y = problems.node_xy(mesh)
    rho_ex, _, _, _ = problems.vortex_fields(x, y, t=20 * dt)
    rho = U1.reshape(s["N_E"], 4, s["N_s"])[:, 0]
    assert np.abs(rho - rho_ex).max() < 1e-5


@pytest.mark.parametrize("c", CONFIGS_PS, ids=cfg_id)
def test_passive_stiffened_limited_trajectories(device, c):
    """Passive-scalar and stiffened-gas Riemann problems with the global HR
    and the sensor-gated HR limiters: 10 CFL-sized steps within 1e-11 of the
    oracle; the sensor flags and one HRI application bit for bit."""
    d, ph, fl = c
    order = 2
    mesh = small_mesh(d, order, ph, n=(24, 10))
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(d, order, ph, fl), mesh)
    U0 = problems.riemann_state(mesh, ph, seed=11) if d == 1 else problems.oblique_riemann_state(mesh, ph, seed=11)
    th = (0.01, 0.01, 0.01)
    solver.set_sensor_thresholds(th)
    Us = _stepped_riemann(mesh, prob, ph, seed=13) if d == 1 else U0 * (1.0 + 0.01 * np.sin(np.arange(U0.size)))
    sens_ref, _ = prob.sensing(Us, th)
    assert sens_ref.any()
    sd = torch.zeros(prob.cfg.N_E, dtype=torch.int32, device=device)
    ua = torch.zeros(prob.cfg.N_E * prob.nf, dtype=torch.float64, device=device)
    Ud = dev(Us, device)
    solver.sensing(Ud, sd, ua)
    assert np.array_equal(host(sd), sens_ref)
    solver.limit(LIMIT_HRI, Ud)
    assert np.array_equal(host(Ud), prob.limit_hri(Us, th)[0])
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / (2 * order + 1))
    for lim in (LIMIT_HR, LIMIT_HRI):
        U_ref, _, _, _ = prob.rk4(U0, 0, dt, nsteps=10, limiter=lim, thresholds=th)
        Ud = dev(U0, device)
        solver.rk4_integrate(Ud, 0, dt, nsteps=10, limiter=lim)
        for l2, li in per_field_rel(host(Ud), U_ref, prob.nf, prob.cfg.N_s):
            assert l2 < 1e-11 and li < 1e-11, (lim, l2, li)


@pytest.mark.parametrize("limiter", [LIMIT_HR, LIMIT_HRI, 8])
@pytest.mark.parametrize("path", ["tiled", "general"])
def test_limited_two_fluid_conserves_totals(device, limiter, path):
    """A periodic two-fluid oblique Riemann problem over 20 limited RK4 steps
    (global HR: the one-launch kernel on the tile path, two sweeps on the
    face + element path; sensor-gated HR; HR on the cells an untrained
    network flags, limiter 8): the totals of rho, rho u, rho v and E stay
    constant to round-off, since the HR limiter keeps every element's mean
    and periodic faces cancel.  (1/(gamma-1) is advected in non-conservative
    form and is not a conserved total.)"""
    from dg_amd import PATH_GENERAL, PATH_TILED, Indicator

    mesh = Mesh(2, 2, 64, 32, 0.0, 2.0, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_path(PATH_TILED if path == "tiled" else PATH_GENERAL)
    solver.set_sensor_thresholds((0.01, 0.01, 0.01))
    if limiter == 8:
        solver.set_indicator(Indicator(seed=3), 0.0)
    if path == "tiled" and limiter == LIMIT_HR:
        assert solver.hr_one_launch
    s = mesh.sizes
    w = mesh.array("weight")
    phi = mesh.array("phi").reshape(s["N_s"], s["N_G"]).T
    J = mesh.array("J")

    def totals(U):
        return np.einsum("g,gi,efi,e->f", w, phi, U.reshape(s["N_E"], 5, s["N_s"]), J)[:4]

    U0 = problems.oblique_riemann_state(mesh, MULTIFLUID, seed=11)
    Ud = dev(U0, device)
    solver.rk4_integrate(Ud, 0, 0.2 * mesh.dx / 5, nsteps=20, limiter=limiter)
    U1 = host(Ud)
    assert not np.array_equal(U1, U0)
    t0, t1 = totals(U0), totals(U1)
    scale = np.abs(np.einsum("g,gi,efi,e->f", w, np.abs(phi), np.abs(U0.reshape(s["N_E"], 5, s["N_s"])), J)[:4])
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale), (t0, t1)
