"""GPU parity of the fused tile kernel (one launch per RK stage on Cartesian meshes).

The tile kernel must give the oracle's right-hand side to the same tolerance
as the face + element kernels (relative 1e-12 on one RHS, 1e-11 on
trajectories), for every physics/flux/order and every boundary combination,
with and without the XCD band mapping of tiles (used when the tile count is
a multiple of 8).
"""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import (FARFIELD, LIMIT_HR, LIMIT_HRI, MULTIFLUID, PATH_GENERAL, PATH_TILED, PERIODIC, REFLECTIVE, ROE,
                    Config, DgxError, Mesh, Solver, nfields, problems)
from tests.helpers import CONFIGS, CONFIGS_PS, cfg_id, per_field_rel

pytestmark = pytest.mark.gpu


def dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


BCS_1D = [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC), (REFLECTIVE, FARFIELD)]
BCS_2D = [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (FARFIELD, REFLECTIVE, PERIODIC, PERIODIC),
          (REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE), (PERIODIC, PERIODIC, PERIODIC, PERIODIC)]
# element counts: tile count not a multiple of 8, then a multiple of 8 (XCD bands)
SIZES_1D = [256, 1024, 200]  # 200: tiles of 8 elements (128 does not divide it)
SIZES_2D = [(16, 8), (32, 16)]


def tiled_mesh(d, order, ph, bc, n):
    nf = nfields(d, ph)
    if d == 1:
        return Mesh(1, order, n, 1, -1.0, 1.0, bc=bc, nfields=nf)
    return Mesh(2, order, n[0], n[1], 0.0, 1.6, 0.0, 0.6, bc=bc, nfields=nf)


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_tiled_rhs_matches_oracle_and_general_path(device, c, order):
    d, ph, fl = c
    for k, bc in enumerate(BCS_1D if d == 1 else BCS_2D):
        n = SIZES_1D[k % 3] if d == 1 else SIZES_2D[k % 2]
        mesh = tiled_mesh(d, order, ph, bc, n)
        prob = oracle.Problem(mesh.tables(), ph, fl)
        solver = Solver(Config(d, order, ph, fl), mesh)
        assert solver.path == PATH_TILED, (bc, n)
        states = [problems.riemann_state(mesh, ph, seed=k + 5)]
        if d == 2 or n <= 256:
            # The RHS of a smooth state is O(h) times its volume and face terms,
            # so rounding relative to max|f| grows with the element count per
            # wavelength; 1024 elements in 1D exceed the 1e-12 bar for either path.
            states.insert(0, problems.smooth_random_state(mesh, ph, seed=k))
        for U in states:
            f_ref = prob.rhs(U)
            Ud = dev(U, device)
            ft = torch.zeros_like(Ud)
            solver.rhs(Ud, ft)
            solver.set_path(PATH_GENERAL)
            fg = torch.zeros_like(Ud)
            solver.rhs(Ud, fg)
            solver.set_path(PATH_TILED)
            scale = np.abs(f_ref).max()
            err_t = np.abs(host(ft) - f_ref).max() / scale
            err_g = np.abs(host(fg) - f_ref).max() / scale
            assert err_t <= 1e-12, (bc, n, err_t)
            assert err_g <= 1e-12, (bc, n, err_g)


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
def test_tiled_rk4_trajectory_with_limiter(device, c):
    d, ph, fl = c
    order = 2
    bc = BCS_1D[0] if d == 1 else BCS_2D[0]
    mesh = tiled_mesh(d, order, ph, bc, SIZES_1D[0] if d == 1 else SIZES_2D[1])
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(d, order, ph, fl), mesh)
    assert solver.path == PATH_TILED
    # 2D: an oblique jump, so that both momenta carry signal (a jump along x
    # leaves rho*v at rounding level, where relative errors mean nothing)
    U0 = problems.riemann_state(mesh, ph, seed=2) if d == 1 else problems.oblique_riemann_state(mesh, ph, seed=2)
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / (2 * order + 1))
    nsteps = 12
    U_ref, n, T, _ = prob.rk4(U0, 0, dt, nsteps=nsteps, limiter=LIMIT_HR)
    Ud = dev(U0, device)
    steps, _ = solver.rk4_integrate(Ud, 0, dt, nsteps=nsteps, limiter=LIMIT_HR)
    assert steps == nsteps
    for l2, li in per_field_rel(host(Ud), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


@pytest.mark.parametrize("ph", [0, MULTIFLUID])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_one_launch_hr_limiter_bitwise(device, ph, order):
    """dgx_limit_into on a tile mesh (both 2D sweeps in one launch, out of
    place) equals the oracle's two sweeps and the in-place limiter bit for
    bit, for every boundary combination, and leaves its input unchanged.
    16 x 12 elements take the 8 x 4 tiles, the other sizes the 8 x 8 ones."""
    cases = [(bc, SIZES_2D[k % 2]) for k, bc in enumerate(BCS_2D)] + [(BCS_2D[0], (16, 12)), (BCS_2D[1], (16, 12))]
    for k, (bc, n) in enumerate(cases):
        mesh = tiled_mesh(2, order, ph, bc, n)
        prob = oracle.Problem(mesh.tables(), ph, ROE)
        solver = Solver(Config(2, order, ph, ROE), mesh)
        assert solver.hr_one_launch == (order <= 2)
        U0 = problems.oblique_riemann_state(mesh, ph, seed=30 + k)
        x = np.arange(U0.size, dtype=np.float64)
        U0 = U0 * (1.0 + 0.01 * np.sin(0.37 * x + k))  # every mode non-trivial
        U_ref = prob.limit_hr(U0)
        assert not np.array_equal(U_ref, U0)
        Ud = dev(U0, device)
        out = torch.full_like(Ud, float("nan"))
        solver.limit_into(LIMIT_HR, Ud, out)
        assert np.array_equal(host(out), U_ref), (bc, n, order)
        assert np.array_equal(host(Ud), U0)
        solver.limit(LIMIT_HR, Ud)
        assert np.array_equal(host(Ud), U_ref)
    with pytest.raises(DgxError):
        solver.limit_into(LIMIT_HR, Ud, Ud)


@pytest.mark.parametrize("c", [x for x in CONFIGS + CONFIGS_PS if x[0] == 1], ids=cfg_id)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_small_1d_tiles_cfl_run_with_hr(device, c, order):
    """1D meshes that 128 does not divide take tiles of 8 elements; with the
    HR limiter the RK driver limits out of place in one launch per stage.
    A CFL Sod run to a final time (76-182 steps) has the oracle's step count
    and time, and every field within 1e-11 of the oracle's state or, where the
    limited shock amplifies rounding beyond that, within twice the largest
    distance between the oracle's run and its own runs from the initial state
    moved by one ulp (all up, all down, and two seeded mixes of both; with
    p = 3 and Rusanov these spread over 2e-11..1.3e-10 in the density's L2)."""
    d, ph, fl = c
    mesh = tiled_mesh(1, order, ph, BCS_1D[0], 200)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    solver = Solver(Config(1, order, ph, fl), mesh)
    assert solver.path == PATH_TILED
    U0 = problems.sod_state(mesh, ph)
    cfl = 0.4 * mesh.dx / (2 * order + 1)
    U_ref, n_ref, T_ref, _ = prob.rk4(U0, 1, cfl, tf=0.05, limiter=LIMIT_HR)
    assert n_ref > 0
    up, down = np.nextafter(U0, np.inf), np.nextafter(U0, -np.inf)
    floor = None
    for U1 in (up, down) + tuple(np.where(np.random.default_rng(k).random(U0.size) < 0.5, up, down) for k in (0, 1)):
        U_ulp, n_ulp, _, _ = prob.rk4(U1, 1, cfl, tf=0.05, limiter=LIMIT_HR)
        assert n_ulp == n_ref
        d = per_field_rel(U_ulp, U_ref, prob.nf, prob.cfg.N_s)
        floor = d if floor is None else [(max(a, c), max(b, e)) for (a, b), (c, e) in zip(floor, d)]
    Ud = dev(U0, device)
    n, T = solver.rk4_integrate(Ud, 1, cfl, tf=0.05, limiter=LIMIT_HR)
    assert n == n_ref and abs(T - T_ref) <= 1e-14
    errs = per_field_rel(host(Ud), U_ref, prob.nf, prob.cfg.N_s)
    for (l2, li), (f2, fi) in zip(errs, floor):
        assert l2 <= max(2 * f2, 1e-11) and li <= max(2 * fi, 1e-11), (errs, floor)


@pytest.mark.parametrize("ph", [0, MULTIFLUID])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_one_launch_hr_limiter_1d_bitwise(device, ph, order):
    """1D dgx_limit_into (one sweep forming its own and the neighbours'
    Taylor coefficients) equals the oracle's HR limiter and the in-place
    two-launch limiter bit for bit, on every boundary combination."""
    for k, bc in enumerate(BCS_1D):
        mesh = tiled_mesh(1, order, ph, bc, 200)
        prob = oracle.Problem(mesh.tables(), ph, ROE)
        solver = Solver(Config(1, order, ph, ROE), mesh)
        U0 = problems.riemann_state(mesh, ph, seed=40 + k)
        x = np.arange(U0.size, dtype=np.float64)
        U0 = U0 * (1.0 + 0.01 * np.sin(0.37 * x + k))
        U_ref = prob.limit_hr(U0)
        assert not np.array_equal(U_ref, U0)
        Ud = dev(U0, device)
        out = torch.full_like(Ud, float("nan"))
        solver.limit_into(LIMIT_HR, Ud, out)
        assert np.array_equal(host(out), U_ref), (bc, order)
        assert np.array_equal(host(Ud), U0)
        solver.limit(LIMIT_HR, Ud)
        assert np.array_equal(host(Ud), U_ref)


def test_tiled_hri_cfl_run(device):
    """Sensor-gated limiting with CFL time steps through the tile kernel."""
    d, ph, fl, order = 2, MULTIFLUID, ROE, 2
    mesh = tiled_mesh(d, order, ph, BCS_2D[1], SIZES_2D[1])
    prob = oracle.Problem(mesh.tables(), ph, fl)
    th = (0.01, 0.01, 0.01)
    solver = Solver(Config(d, order, ph, fl), mesh)
    solver.set_sensor_thresholds(th)
    assert solver.path == PATH_TILED
    U0 = problems.oblique_riemann_state(mesh, ph, seed=4)
    cfl = 0.3 * mesh.dx / (2 * order + 1)
    U_ref, n_ref, T_ref, dts = prob.rk4(U0, 1, cfl, tf=0.02, limiter=LIMIT_HRI, thresholds=th)
    Ud = dev(U0, device)
    n, T = solver.rk4_integrate(Ud, 1, cfl, tf=0.02, limiter=LIMIT_HRI)
    assert n == n_ref and abs(T - T_ref) <= 1e-12 * T_ref
    for l2, li in per_field_rel(host(Ud), U_ref, prob.nf, prob.cfg.N_s):
        assert l2 < 1e-11 and li < 1e-11, (l2, li)


def test_path_selection(device):
    ph = MULTIFLUID
    # 12 x 7 elements is not a multiple of the 8 x 4 tile: general path only
    mesh = tiled_mesh(2, 2, ph, BCS_2D[0], (12, 7))
    solver = Solver(Config(2, 2, ph, ROE), mesh)
    assert solver.path == PATH_GENERAL
    with pytest.raises(DgxError):
        solver.set_path(PATH_TILED)
    mesh = tiled_mesh(2, 2, ph, BCS_2D[0], (16, 8))
    solver = Solver(Config(2, 2, ph, ROE), mesh)
    assert solver.path == PATH_TILED
    solver.set_path(PATH_GENERAL)
    assert solver.path == PATH_GENERAL
    mesh = tiled_mesh(1, 2, ph, BCS_1D[0], 203)
    assert Solver(Config(1, 2, ph, ROE), mesh).path == PATH_GENERAL
    mesh = tiled_mesh(1, 2, ph, BCS_1D[0], 200)
    assert Solver(Config(1, 2, ph, ROE), mesh).path == PATH_TILED


def test_one_launch_hr_full_size_equals_two_sweeps(device):
    """At the shock-bubble's full size (1024 x 512, p = 2, two-fluid, walls
    above and below), the one-launch limiter (dgx_limit_into) and the
    in-place two-sweep limiter (dgx_limit) give the same bits."""
    mesh = problems.shock_bubble_mesh(1024, 512, 2)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    assert solver.hr_one_launch
    U = dev(problems.shock_bubble_state(mesh), device)
    # a few unlimited steps, so that every mode of the state is non-trivial
    solver.rk4_integrate(U, 0, 2e-4, nsteps=3)
    out = torch.full_like(U, float("nan"))
    solver.limit_into(LIMIT_HR, U, out)
    solver.limit(LIMIT_HR, U)
    torch.cuda.synchronize()
    assert torch.equal(out, U)


def test_one_launch_hr_output_not_16_byte_aligned(device):
    """The one-launch limiter stores pairs of doubles only into a 16-byte
    aligned output; an output (or an RK state) that starts 8 bytes past such a
    boundary takes single stores and the same bits."""
    mesh = tiled_mesh(2, 2, MULTIFLUID, BCS_2D[0], (16, 8))
    prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    assert solver.hr_one_launch
    U0 = problems.oblique_riemann_state(mesh, MULTIFLUID, seed=7)
    U_ref = prob.limit_hr(U0)
    Ud = dev(U0, device)
    big = torch.full((U0.size + 1,), float("nan"), dtype=torch.float64, device=device)
    out = big[1:]
    assert out.data_ptr() % 16 == 8
    solver.limit_into(LIMIT_HR, Ud, out)
    assert np.array_equal(host(out), U_ref)
    # the RK driver's last limiter application writes the caller's state
    dt = 0.3 * prob.dt_from_cfl(U0, mesh.dx / 5)
    Ua = dev(U0, device)
    solver.rk4_integrate(Ua, 0, dt, nsteps=3, limiter=LIMIT_HR)
    big2 = torch.zeros(U0.size + 1, dtype=torch.float64, device=device)
    Ub = big2[1:]
    Ub.copy_(dev(U0, device))
    solver.rk4_integrate(Ub, 0, dt, nsteps=3, limiter=LIMIT_HR)
    assert torch.equal(Ua, Ub)
