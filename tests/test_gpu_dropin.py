"""GPU tests of the reference-named drop-in launchers (include/dgx.h group 1).

The reference's own DG_SOLVER::dg_solver call sequence (dg_solver.h:292-364)
is replayed on device buffers through the drop-in launchers and the BLAS
replacement, and compared with the CPU oracle; the limiter/sensor launchers
must be bit-identical to the oracle (and hence to the reference).
"""
import ctypes

import numpy as np
import pytest
import torch

import dg_amd
import oracle
from dg_amd import MULTIFLUID, PASSIVE, ROE, RUSANOV, SINGLEFLUID, STIFFENED, Config, problems
from tests.helpers import cfg_id, small_mesh

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def L():
    lib = dg_amd.lib()
    i, d, b = ctypes.c_int, ctypes.c_double, ctypes.c_bool
    sig = {
        "LmapToFace": [i, i, i, vp, vp, vp], "LmapToElement": [i, i, i, i, vp, vp, vp],
        "Lredistribute_sf": [i, i, vp, vp, vp, vp, vp, vp], "Lredistribute_q": [i, i, vp, vp, vp],
        "LaddSFQ": [i, i, vp, vp, vp, vp], "Levaluate_sf": [i, i, vp, vp, vp, vp, vp],
        "Levaluate_q": [i, i, vp, vp, vp], "LrflctiveBoundary": [i, i, vp, vp, i, vp],
        "Lsolver": [i, i, d, vp, vp, vp], "LfindUPA": [i, i, vp, vp],
        "Lhrl1D": [i, i, i, i, i, vp, i, vp, vp], "Lhri1D": [i, i, i, vp, i, i, i, vp, vp, vp, vp, vp],
        "Lcalc_sensors": [i, i, b, d, b, d, b, d, vp, vp, vp], "Lcopy_detected": [i, i, vp, vp, vp],
        "dgx_blasDgemm": [ctypes.c_char, ctypes.c_char, i, i, i, d, vp, i, vp, i, d, vp, i],
        "dgx_blasDaxpy": [i, d, vp, i, vp, i], "dgx_blasDcopy": [i, vp, i, vp, i], "dgx_blasIdamax": [i, vp, i],
        "Lstridedcopy": [i, i, i, i, i, i, vp, vp], "Lreconstruct_energy": [i, i, i, vp, vp, vp, vp],
        "Linternal_energy_multifluid": [i, i, i, vp, vp, vp], "Linternal_energy_stiffened": [i, i, i, vp, vp, vp, vp],
        "Lm2i1D": [i, i, i, vp, i, i, i, vp, vp, vp, vp, vp], "Lp0i": [i, i, i, d, vp, vp, vp, vp],
        "LChangeBasis": [i, i, i, vp, vp, vp], "LPrim2Cons": [i, i, vp], "LCons2Prim": [i, i, vp],
        "Lpressure": [i, i, vp, vp], "Lkinetic_energy1D": [i, i, vp, vp, vp], "Lkinetic_energy2D": [i, i, vp, vp, vp, vp],
        "Laverage_cell_p0": [i, i, vp], "Lformater": [i, i, vp, vp, b], "Lformat_sensor": [i, i, vp, vp],
    }
    for n, a in sig.items():
        getattr(lib, n).argtypes = a
    return lib


def D(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


class _DevPtr(vp):
    """A device pointer that keeps its tensor alive for the call it is passed
    to (a temporary's block could otherwise be handed to the next argument)."""


def P(t):
    p = _DevPtr(t.data_ptr())
    p.keep = t
    return p


def H(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [(1, MULTIFLUID, RUSANOV), (1, SINGLEFLUID, ROE), (2, SINGLEFLUID, ROE), (2, MULTIFLUID, ROE),
         (1, PASSIVE, ROE), (2, PASSIVE, RUSANOV), (1, STIFFENED, RUSANOV), (2, STIFFENED, ROE)]


@pytest.mark.parametrize("c", CASES, ids=cfg_id)
def test_dg_solver_sequence_with_dropin_launchers(device, c):
    d, ph, fl = c
    lib = L()
    mesh = small_mesh(d, 2, ph)
    t = mesh.tables()
    prob = oracle.Problem(t, ph, fl)
    cfg = Config(d, 2, ph, fl)
    assert lib.dgx_set_config(ctypes.byref(cfg)) == 0
    U = problems.riemann_state(mesh, ph, seed=3) * (1 + 0.01 * np.sin(np.arange(prob.state_size)))
    NF, Ns, NG, Ms, MG, NE, MT = prob.nf, t["N_s"], t["N_G"], t["M_s"], t["M_G"], t["N_E"], t["M_T"]
    g = {k: D(t[k], device) for k in ("map", "invmap", "boundaryMap", "phi", "dphi", "psi", "phi_w", "dphi_w",
                                        "psi_w", "J", "invJac", "JF", "normals", "Minv")}
    z = lambda n: torch.zeros(n, dtype=torch.float64, device=device)
    Ud = D(U, device)
    UF, Ug, dUg, UgF = z(2 * NF * Ms * MT), z(NF * NG * NE), z(d * NG * NE * NF), z(2 * NF * MG * MT)
    s, sJ, S, f, fJ, F = z(NG * NE * NF), z(NG * NE * NF), z(Ns * NE * NF), z(d * NF * NG * NE), z(d * NG * NE * NF), z(Ns * NE * NF)
    q, qJ, Qt, Q, frk = z(MG * MT * NF * 2), z(MG * MT * NF * 2), z(Ms * MT * NF * 2), z(Ns * NE * NF), z(Ns * NE * NF)
    lib.LmapToFace(Ms, MT, Ns, P(g["map"]), P(Ud), P(UF))
    lib.LrflctiveBoundary(Ms, t["rflctiveIdx"], P(g["boundaryMap"]), P(g["normals"]), 0, P(UF))
    lib.dgx_blasDgemm(b"N", b"N", NG, NE * NF, Ns, 1.0, P(g["phi"]), NG, P(Ud), Ns, 0.0, P(Ug), NG)
    lib.dgx_blasDgemm(b"N", b"N", NG * d, NE * NF, Ns, 1.0, P(g["dphi"]), NG * d, P(Ud), Ns, 0.0, P(dUg), NG * d)
    lib.dgx_blasDgemm(b"N", b"N", MG, MT * NF * 2, Ms, 1.0, P(g["psi"]), MG, P(UF), Ms, 0.0, P(UgF), MG)
    lib.Levaluate_sf(NG, NE, P(s), P(f), P(Ug), P(dUg), P(g["invJac"]))
    lib.Levaluate_q(MG, MT, P(q), P(UgF), P(g["normals"]))
    lib.Lredistribute_sf(NG, NE, P(sJ), P(fJ), P(s), P(f), P(g["J"]), P(g["invJac"]))
    lib.dgx_blasDgemm(b"T", b"N", Ns, NE * NF, NG, 1.0, P(g["phi_w"]), NG, P(sJ), NG, 0.0, P(S), Ns)
    lib.dgx_blasDgemm(b"T", b"N", Ns, NE * NF, NG * d, 1.0, P(g["dphi_w"]), NG * d, P(fJ), NG * d, 0.0, P(F), Ns)
    lib.Lredistribute_q(MG, MT, P(qJ), P(q), P(g["JF"]))
    lib.dgx_blasDgemm(b"T", b"N", Ms, MT * NF * 2, MG, 1.0, P(g["psi_w"]), MG, P(qJ), MG, 0.0, P(Qt), Ms)
    lib.LmapToElement(Ns, NE, Ms, t["N_N"], P(g["invmap"]), P(Q), P(Qt))
    lib.LaddSFQ(Ns, NE, P(frk), P(S), P(F), P(Q))
    f_ref = prob.rhs(U)
    f_gpu = H(frk)
    assert np.abs(f_gpu - f_ref).max() <= 1e-12 * np.abs(f_ref).max()
    # solve + findUPA + blas helpers
    du = z(prob.state_size)
    lib.Lsolver(Ns, NE, 0.013, P(g["Minv"]), P(frk), P(du))
    assert np.abs(H(du) - prob.solve(0.013, f_ref)).max() <= 1e-12 * np.abs(prob.solve(0.013, f_ref)).max()
    upa = z(NE * Ns)
    lib.LfindUPA(Ns, NE, P(Ud), P(upa))
    u_h = H(upa)
    assert prob.dt_from_cfl(U, 1.0) == 1.0 / u_h.max()
    assert lib.dgx_blasIdamax(NE * Ns, P(upa), 1) == int(np.argmax(np.abs(u_h))) + 1
    y = D(U, device)
    lib.dgx_blasDaxpy(prob.state_size, 0.5, P(du), 1, P(y), 1)
    assert np.array_equal(H(y), U + 0.5 * H(du))
    lib.dgx_blasDcopy(prob.state_size, P(Ud), 1, P(y), 1)
    assert np.array_equal(H(y), U)


@pytest.mark.parametrize("dim", [1, 2])
def test_limiter_and_sensor_launchers_bitwise(device, dim):
    ph, fl = MULTIFLUID, ROE
    lib = L()
    mesh = small_mesh(dim, 2, ph)
    t = mesh.tables()
    prob = oracle.Problem(t, ph, fl)
    cfg = Config(dim, 2, ph, fl)
    assert lib.dgx_set_config(ctypes.byref(cfg)) == 0
    U = problems.riemann_state(mesh, ph, seed=8) * (1 + 0.01 * np.sin(np.arange(prob.state_size)))
    NF, Ns, NE, NN = prob.nf, t["N_s"], t["N_E"], t["N_N"]
    nb = D(t["neighbors"], device)
    # HR slices on an arbitrary monomial array
    rng = np.random.default_rng(2)
    A = rng.standard_normal(prob.state_size)
    n1 = 3
    args = [(Ns, 1, 0)] if dim == 1 else [(n1, n1, 0), (n1, n1, 2)]
    for ns, sl, off in args:
        a_ref = np.zeros_like(A)
        oracle.lib().or_hrl1D(ns, NE, NF, NN, sl, t["neighbors"].ctypes.data_as(vp), off, A.ctypes.data_as(vp),
                              a_ref.ctypes.data_as(vp))
        Ad = D(A, device)
        out = torch.zeros_like(Ad)
        lib.Lhrl1D(ns, NE, NF, NN, sl, P(nb), off, P(Ad), P(out))
        assert np.array_equal(H(out), a_ref)
    # sensors (sequential-order semantics) from the reference's Uavg
    th = (0.01, 0.01, 0.01)
    sens_ref, uavg = prob.sensing(U, th)
    sd = torch.zeros(NE, dtype=torch.int32, device=device)
    lib.Lcalc_sensors(NE, NN, True, th[0], True, th[1], True, th[2], P(nb), P(D(uavg, device)), P(sd))
    assert np.array_equal(H(sd), sens_ref)
    # hri1D + copy_detected (x sweep)
    L2M = D(t["Lag2Mono" if dim == 1 else "Lag2MonoX"], device)
    M2L = D(t["Mono2Lag" if dim == 1 else "MonoX2Lag"], device)
    Ud = D(U, device)
    Unew = torch.zeros_like(Ud)
    lib.Lhri1D(Ns, NE, NN, P(nb), n1 if dim == 2 else Ns, n1 if dim == 2 else 1, 0, P(L2M), P(M2L), P(sd), P(Ud),
               P(Unew))
    lib.Lcopy_detected(Ns, NE, P(sd), P(Unew), P(Ud))
    if dim == 2:  # y sweep (limiting.cc:627-628)
        L2M, M2L = D(t["Lag2MonoY"], device), D(t["MonoY2Lag"], device)
        lib.Lhri1D(Ns, NE, NN, P(nb), n1, n1, 2, P(L2M), P(M2L), P(sd), P(Ud), P(Unew))
        lib.Lcopy_detected(Ns, NE, P(sd), P(Unew), P(Ud))
    U_ref, s2 = prob.limit_hri(U, th)
    assert np.array_equal(s2, sens_ref)
    assert np.array_equal(H(Ud), U_ref)
    assert not np.array_equal(U_ref, U)
    assert (sens_ref != 0).any()


def _cfg(lib, c, order=2):
    assert lib.dgx_set_config(ctypes.byref(Config(c[0], order, c[1], c[2]))) == 0


def _vp(a):
    return a.ctypes.data_as(vp)


HELPER_CASES = [(1, SINGLEFLUID, ROE), (2, SINGLEFLUID, ROE), (1, MULTIFLUID, ROE), (2, MULTIFLUID, RUSANOV),
                (1, PASSIVE, ROE), (2, PASSIVE, ROE), (1, STIFFENED, ROE), (2, STIFFENED, RUSANOV)]


@pytest.mark.parametrize("c", HELPER_CASES, ids=cfg_id)
def test_helper_launchers_bitwise(device, c):
    """Lpressure, Lkinetic_energy1D/2D, Laverage_cell_p0, LChangeBasis,
    LCons2Prim/LPrim2Cons, Lformater, Lformat_sensor against the oracle's
    restatements (which equal the reference's launchers bit for bit)."""
    d, ph, fl = c
    lib, O = L(), oracle.lib()
    _cfg(lib, c, order=3)
    mesh = small_mesh(d, 3, ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    cfgp = ctypes.byref(prob.cfg)
    NE, Ns, NF = prob.cfg.N_E, prob.cfg.N_s, prob.nf
    U = problems.riemann_state(mesh, ph, seed=11)
    Ud = D(U, device)
    p_ref = np.zeros(NE * Ns)
    O.or_pressure(cfgp, _vp(U), _vp(p_ref))
    pd = torch.zeros(NE * Ns, dtype=torch.float64, device=device)
    lib.Lpressure(Ns, NE, P(Ud), P(pd))
    assert np.array_equal(H(pd), p_ref)
    V = U.reshape(NE, NF, Ns)
    rho, ru = np.ascontiguousarray(V[:, 0]).ravel(), np.ascontiguousarray(V[:, 1]).ravel()
    rv = np.ascontiguousarray(V[:, 2]).ravel() if d == 2 else None
    k_ref = np.zeros(NE * Ns)
    O.or_kinetic_energy(Ns, NE, _vp(rho), _vp(ru), _vp(rv) if d == 2 else None, _vp(k_ref))
    kd = torch.zeros_like(pd)
    if d == 1:
        lib.Lkinetic_energy1D(Ns, NE, P(D(rho, device)), P(D(ru, device)), P(kd))
    else:
        lib.Lkinetic_energy2D(Ns, NE, P(D(rho, device)), P(D(ru, device)), P(D(rv, device)), P(kd))
    assert np.array_equal(H(kd), k_ref)
    rng = np.random.default_rng(5)
    DU = rng.standard_normal(U.size)
    DUd = D(DU, device)
    O.or_average_cell_p0(cfgp, _vp(DU))
    lib.Laverage_cell_p0(Ns, NE, P(DUd))
    assert np.array_equal(H(DUd), DU)
    for s1, s2 in ((Ns, Ns), (Ns - 1, Ns)):
        T = rng.standard_normal(NE * s1 * s2)
        X = rng.standard_normal(NE * NF * s2)
        o_ref = np.zeros(NE * NF * s1)
        O.or_change_basis(cfgp, s1, s2, _vp(T), _vp(X), _vp(o_ref))
        od = torch.zeros(NE * NF * s1, dtype=torch.float64, device=device)
        lib.LChangeBasis(s1, s2, NE, P(D(T, device)), P(D(X, device)), P(od))
        assert np.array_equal(H(od), o_ref)
    if ph != SINGLEFLUID:
        W = U.copy()
        O.or_cons_prim(cfgp, _vp(W), 1)
        Wd = D(U, device)
        lib.LCons2Prim(Ns, NE, P(Wd))
        assert np.array_equal(H(Wd), W)
        O.or_cons_prim(cfgp, _vp(W), 0)
        lib.LPrim2Cons(Ns, NE, P(Wd))
        assert np.array_equal(H(Wd), W)
    if ph != PASSIVE:  # the formatter (Lformater) both ways, as PRINTER::print / read use it
        out = torch.zeros_like(Ud)
        lib.Lformater(Ns, NE, P(Ud), P(out), False)
        assert np.array_equal(H(out), prob.format(U))
        back = torch.zeros_like(Ud)
        lib.Lformater(Ns, NE, P(back), P(out), True)
        assert np.array_equal(H(back), prob.unformat(prob.format(U)))
    sens = rng.integers(0, 4, NE).astype(np.int32)
    fs_ref = np.zeros(NE * Ns)
    O.or_format_sensor(Ns, NE, _vp(sens), _vp(fs_ref))
    fs = torch.zeros(NE * Ns, dtype=torch.float64, device=device)
    lib.Lformat_sensor(Ns, NE, P(D(sens, device)), P(fs))
    assert np.array_equal(H(fs), fs_ref)


def test_energy_and_copy_launchers_bitwise(device):
    lib, O = L(), oracle.lib()
    _cfg(lib, (2, STIFFENED, ROE))
    rng = np.random.default_rng(6)
    A = rng.standard_normal(600)
    for args in ((2, 3, 9, 3, 0, 0), (7, 4, 20, 5, 3, 1), (12, 9, 45, 9, 18, 2)):
        nb, bl, sa, sb, oa, ob = args
        assert oa + (nb - 1) * sa + bl <= A.size and ob + (nb - 1) * sb + bl <= 300  # every copied block in bounds
        B_ref = np.full(300, 7.0)
        O.or_stridedcopy(*args, _vp(A), _vp(B_ref))
        Bd = D(np.full(300, 7.0), device)
        lib.Lstridedcopy(*args, P(D(A, device)), P(Bd))
        assert np.array_equal(H(Bd), B_ref)
    NE = 37
    for Ns, sl in ((3, 1), (3, 3), (4, 4)):
        n = Ns * sl * NE
        p, g, b, K, E = (rng.standard_normal(n) for _ in range(5))
        for bb in (None, b):
            r_ref = np.zeros(n)
            O.or_internal_energy(Ns, NE, sl, _vp(p), _vp(g), _vp(bb) if bb is not None else None, _vp(r_ref))
            rd = torch.zeros(n, dtype=torch.float64, device=device)
            if bb is None:
                lib.Linternal_energy_multifluid(Ns, NE, sl, P(D(p, device)), P(D(g, device)), P(rd))
            else:
                lib.Linternal_energy_stiffened(Ns, NE, sl, P(D(p, device)), P(D(g, device)), P(D(bb, device)), P(rd))
            assert np.array_equal(H(rd), r_ref)
        e_ref = np.zeros(n)
        O.or_reconstruct_energy(Ns, NE, sl, _vp(r_ref), _vp(K), _vp(E), _vp(e_ref))
        ed = torch.zeros(n, dtype=torch.float64, device=device)
        lib.Lreconstruct_energy(Ns, NE, sl, P(D(r_ref, device)), P(D(K, device)), P(D(E, device)), P(ed))
        assert np.array_equal(H(ed), e_ref)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2])
def test_m2_limiting_replay_bitwise(device, dim, order):
    """Limiting::M2limiting (limiting.cc:123-594, two-fluid) replayed through
    the launchers in the reference's call order (Lstridedcopy, dgx_blasDgemm,
    Lhrl1D, Lpressure, Lkinetic_energy*, Linternal_energy_multifluid,
    Lreconstruct_energy) equals the oracle's or_limit_m2 bit for bit."""
    from dg_amd import PERIODIC, REFLECTIVE, FARFIELD

    ph, fl = MULTIFLUID, ROE
    lib = L()
    _cfg(lib, (dim, ph, fl), order)
    bc = (FARFIELD, REFLECTIVE) if dim == 1 else (PERIODIC, PERIODIC, REFLECTIVE, FARFIELD)
    mesh = small_mesh(dim, order, ph, bc=bc)
    t = mesh.tables()
    prob = oracle.Problem(t, ph, fl)
    NE, Ns, NF, NN, n1 = prob.cfg.N_E, prob.cfg.N_s, prob.nf, prob.cfg.N_N, prob.cfg.N_s1D
    U = problems.riemann_state(mesh, ph, seed=31) * (1 + 0.01 * np.sin(0.37 * np.arange(prob.state_size)))
    Ud = D(U, device)
    nb = D(t["neighbors"], device)
    T = {k: D(t[k], device) for k in ("Lag2Mono", "Mono2Lag", "Lag2MonoX", "MonoX2Lag", "Lag2MonoY", "MonoY2Lag")
         if t[k].size}
    z = lambda: torch.zeros(Ns * NE, dtype=torch.float64, device=device)
    rho, rhou, rhov, E, gam, pres, K, mono, lim, gLim, pLim, KLim, EMono, rhoeLim, ELim = (z() for _ in range(15))
    gemm = lambda M, X, Y: lib.dgx_blasDgemm(b"N", b"N", Ns, NE, Ns, 1.0, P(M), Ns, P(X), Ns, 0.0, P(Y), Ns)
    get = lambda f, x: lib.Lstridedcopy(NE, Ns, Ns * NF, Ns, f * Ns, 0, P(Ud), P(x))
    put = lambda f, x: lib.Lstridedcopy(NE, Ns, Ns, Ns * NF, 0, f * Ns, P(x), P(Ud))
    if dim == 1:
        def hr(x, Lm, back):
            gemm(T["Lag2Mono"], x, mono)
            lib.Lhrl1D(Ns, NE, 1, NN, 1, P(nb), 0, P(mono), P(Lm))
            if back:
                gemm(T["Mono2Lag"], Lm, x)
        get(0, rho); hr(rho, lim, True)
        get(1, rhou); hr(rhou, lim, True)
        get(2, E); gemm(T["Lag2Mono"], E, EMono)
        get(3, gam); hr(gam, gLim, True)
        lib.Lpressure(Ns, NE, P(Ud), P(pres)); hr(pres, pLim, False)
        lib.Lkinetic_energy1D(Ns, NE, P(rho), P(rhou), P(K)); gemm(T["Lag2Mono"], K, KLim)
        lib.Linternal_energy_multifluid(Ns, NE, 1, P(pLim), P(gLim), P(rhoeLim))
        lib.Lreconstruct_energy(Ns, NE, 1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
        gemm(T["Mono2Lag"], ELim, E)
        for f, x in enumerate((rho, rhou, E, gam)):
            put(f, x)
    else:
        def hr(L2M, M2L, off, x, Lm, back):
            gemm(T[L2M], x, mono)
            lib.Lhrl1D(n1, NE, 1, NN, n1, P(nb), off, P(mono), P(Lm))
            if back:
                gemm(T[M2L], Lm, x)
        for f, x in ((0, rho), (1, rhou), (2, rhov)):
            get(f, x); hr("Lag2MonoX", "MonoX2Lag", 0, x, lim, True)
        get(4, gam); hr("Lag2MonoX", "MonoX2Lag", 0, gam, gLim, True)
        lib.Lpressure(Ns, NE, P(Ud), P(pres)); hr("Lag2MonoX", "MonoX2Lag", 0, pres, pLim, True)
        lib.Lkinetic_energy2D(Ns, NE, P(rho), P(rhou), P(rhov), P(K)); gemm(T["Lag2MonoX"], K, KLim)
        lib.Linternal_energy_multifluid(n1, NE, n1, P(pLim), P(gLim), P(rhoeLim))
        get(3, E); gemm(T["Lag2MonoX"], E, EMono)
        lib.Lreconstruct_energy(n1, NE, n1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
        gemm(T["MonoX2Lag"], ELim, E)
        for x in (rho, rhou, rhov):
            hr("Lag2MonoY", "MonoY2Lag", 2, x, lim, True)
        hr("Lag2MonoY", "MonoY2Lag", 2, gam, gLim, True)
        hr("Lag2MonoY", "MonoY2Lag", 2, pres, pLim, False)
        lib.Lkinetic_energy2D(Ns, NE, P(rho), P(rhou), P(rhov), P(K)); gemm(T["Lag2MonoY"], K, KLim)
        lib.Linternal_energy_multifluid(n1, NE, n1, P(pLim), P(gLim), P(rhoeLim))
        gemm(T["Lag2MonoY"], E, EMono)
        lib.Lreconstruct_energy(n1, NE, n1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
        gemm(T["MonoY2Lag"], ELim, E)
        for f, x in enumerate((rho, rhou, rhov, E, gam)):
            put(f, x)
    U_ref = prob.limit_m2(U)
    assert not np.array_equal(U_ref, U)
    assert np.array_equal(H(Ud), U_ref)


@pytest.mark.parametrize("dim", [1, 2])
def test_sensor_gated_launchers_replay_bitwise(device, dim):
    """P0Ilimiting (limiting.cc:723-744: sensing, Lp0i, Lcopy_detected),
    M2Ilimiting (limiting.cc:682-720: Lm2i1D + Lcopy_detected per direction)
    and PRlimiting (limiting.cc:102-121: LCons2Prim, HRlimiting through
    dgx_blasDgemm and Lhrl1D, LPrim2Cons) through the launchers, against
    the oracle, bit for bit."""
    from dg_amd import PERIODIC, REFLECTIVE, FARFIELD

    ph, fl = MULTIFLUID, ROE
    lib = L()
    _cfg(lib, (dim, ph, fl))
    bc = (FARFIELD, REFLECTIVE) if dim == 1 else (PERIODIC, PERIODIC, REFLECTIVE, FARFIELD)
    mesh = small_mesh(dim, 2, ph, bc=bc)
    t = mesh.tables()
    prob = oracle.Problem(t, ph, fl)
    NE, Ns, NN, NG, n1 = prob.cfg.N_E, prob.cfg.N_s, prob.cfg.N_N, prob.cfg.N_G, prob.cfg.N_s1D
    # a milder ripple in 1D, where 0.01 would make the sensor flag every cell
    amp = 0.01 if dim == 2 else 0.001
    U = problems.riemann_state(mesh, ph, seed=41) * (1 + amp * np.sin(0.41 * np.arange(prob.state_size)))
    th = (0.01, 0.01, 0.01)
    sens, uavg = prob.sensing(U, th)
    assert sens.any() and not sens.all()
    nb, sd = D(t["neighbors"], device), D(sens, device)
    # P0I
    Ud = D(U, device)
    Un = torch.zeros_like(Ud)
    area = 2.0 if dim == 1 else 4.0
    lib.Lp0i(Ns, NE, NG, area, P(sd), P(D(t["phi_w"], device)), P(Ud), P(Un))
    lib.Lcopy_detected(Ns, NE, P(sd), P(Un), P(Ud))
    W, s2 = prob.limit_p0i(U, th)
    assert np.array_equal(s2, sens) and np.array_equal(H(Ud), W)
    # M2I
    Ud = D(U, device)
    sweeps = ([("Lag2Mono", "Mono2Lag", Ns, 1, 0)] if dim == 1 else
              [("Lag2MonoX", "MonoX2Lag", n1, n1, 0), ("Lag2MonoY", "MonoY2Lag", n1, n1, 2)])
    for l2m, m2l, ns1, sl, off in sweeps:
        Un = torch.zeros_like(Ud)
        lib.Lm2i1D(Ns, NE, NN, P(nb), ns1, sl, off, P(D(t[l2m], device)), P(D(t[m2l], device)), P(sd), P(Ud), P(Un))
        lib.Lcopy_detected(Ns, NE, P(sd), P(Un), P(Ud))
    W, _ = prob.limit_m2i(U, th)
    assert not np.array_equal(W, U) and np.array_equal(H(Ud), W)
    # PR
    Ud = D(U, device)
    lib.LCons2Prim(Ns, NE, P(Ud))
    A, Alim = torch.zeros_like(Ud), torch.zeros_like(Ud)
    NF = prob.nf
    gemm = lambda M, X, Y: lib.dgx_blasDgemm(b"N", b"N", Ns, NE * NF, Ns, 1.0, P(M), Ns, P(X), Ns, 0.0, P(Y), Ns)
    if dim == 1:
        gemm(D(t["Lag2Mono"], device), Ud, A)
        lib.Lhrl1D(Ns, NE, NF, NN, 1, P(nb), 0, P(A), P(Alim))
        gemm(D(t["Mono2Lag"], device), Alim, Ud)
    else:
        gemm(D(t["Lag2MonoX"], device), Ud, A)
        lib.Lhrl1D(n1, NE, NF, NN, n1, P(nb), 0, P(A), P(Alim))
        gemm(D(t["MonoX2MonoY"], device), Alim, A)
        lib.Lhrl1D(n1, NE, NF, NN, n1, P(nb), 2, P(A), P(Alim))
        gemm(D(t["MonoY2Lag"], device), Alim, Ud)
    lib.LPrim2Cons(Ns, NE, P(Ud))
    assert np.array_equal(H(Ud), prob.limit_pr(U))


def test_idamax_passes_over_nan_like_blas(device):
    """dgx_blasIdamax follows the reference BLAS: the first largest |x|, a NaN
    after x[0] passed over, a NaN x[0] returned (1-based)."""
    lib = L()
    x = np.array([0.5, -3.0, np.nan, 3.0, -2.0, np.nan])
    assert lib.dgx_blasIdamax(x.size, P(D(x, device)), 1) == 2
    x[0] = np.nan
    assert lib.dgx_blasIdamax(x.size, P(D(x, device)), 1) == 1
    y = np.array([0.0, 0.0, np.nan])
    assert lib.dgx_blasIdamax(y.size, P(D(y, device)), 1) == 1
