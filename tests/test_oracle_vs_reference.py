"""Pin the CPU oracle to the reference's own right-hand-side kernel files.

The reference's launchers (dg/kernels.cu, physics/physics.cu,
time_integration/rk_kernels.cu, limiting/limiting_kernels.cu,
sensors/sensor_kernels.cu, bc/boundaries.cu) are compiled from the reference
sources as CPU loops (ARCH=USE_CPU), one library per physics configuration
(oracle/Makefile -> oracle/_ref/).  Every oracle component must equal the
reference's output bit for bit on the same inputs, and the oracle's whole
right-hand side must equal the reference launcher sequence of
DG_SOLVER::dg_solver (dg_solver.h:292-364) composed with netlib-order GEMMs.
"""
import ctypes

import numpy as np
import pytest

import oracle
from dg_amd import FARFIELD, MULTIFLUID, PERIODIC, REFLECTIVE, STIFFENED, problems
from tests.helpers import CONFIGS, CONFIGS_PS, cfg_id, small_mesh
from tests.test_oracle_compositions import ref_hr, ref_hri_sweeps

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="reference sources not present: _ref not built")

vp = ctypes.c_void_p


def P(a):
    return a.ctypes.data_as(vp)


def _setup(c, order, seed, riemann=False, bc=None):
    d, ph, fl = c
    mesh = small_mesh(d, order, ph, bc=bc)
    t = mesh.tables()
    prob = oracle.Problem(t, ph, fl)
    U = (problems.riemann_state(mesh, ph, seed) if riemann else problems.smooth_random_state(mesh, ph, seed))
    return mesh, t, prob, U


def _ref_rhs(R, prob, t, U):
    """DG_SOLVER::dg_solver with the reference launchers + netlib dgemm."""
    c = prob.cfg
    L = oracle.lib()
    NF, Ns, NG, Ms, MG, NE, MT, D = prob.nf, c.N_s, c.N_G, c.M_s, c.M_G, c.N_E, c.M_T, c.D
    z = lambda n: np.zeros(n)
    UF, Ug, dUg, UgF = z(2 * NF * Ms * MT), z(NF * NG * NE), z(D * NG * NE * NF), z(2 * NF * MG * MT)
    s, sJ, S, f, fJ, F = z(NG * NE * NF), z(NG * NE * NF), z(Ns * NE * NF), z(D * NF * NG * NE), z(D * NG * NE * NF), z(Ns * NE * NF)
    q, qJ, Qt, Q, frk = z(MG * MT * NF * 2), z(MG * MT * NF * 2), z(Ms * MT * NF * 2), z(Ns * NE * NF), z(Ns * NE * NF)
    k = prob._keep
    R.LmapToFace(Ms, MT, Ns, P(k["map"]), P(U), P(UF))
    R.LrflctiveBoundary(Ms, int(t["rflctiveIdx"]), P(k["boundaryMap"]), P(k["normals"]), 0, P(UF))
    L.or_dgemm(b"N", b"N", NG, NE * NF, Ns, 1.0, P(k["phi"]), NG, P(U), Ns, 0.0, P(Ug), NG)
    L.or_dgemm(b"N", b"N", NG * D, NE * NF, Ns, 1.0, P(k["dphi"]), NG * D, P(U), Ns, 0.0, P(dUg), NG * D)
    L.or_dgemm(b"N", b"N", MG, MT * NF * 2, Ms, 1.0, P(k["psi"]), MG, P(UF), Ms, 0.0, P(UgF), MG)
    R.Levaluate_sf(NG, NE, P(s), P(f), P(Ug), P(dUg), P(k["invJac"]))
    R.Levaluate_q(MG, MT, P(q), P(UgF), P(k["normals"]))
    R.Lredistribute_sf(NG, NE, P(sJ), P(fJ), P(s), P(f), P(k["J"]), P(k["invJac"]))
    L.or_dgemm(b"T", b"N", Ns, NE * NF, NG, 1.0, P(k["phi_w"]), NG, P(sJ), NG, 0.0, P(S), Ns)
    L.or_dgemm(b"T", b"N", Ns, NE * NF, NG * D, 1.0, P(k["dphi_w"]), NG * D, P(fJ), NG * D, 0.0, P(F), Ns)
    R.Lredistribute_q(MG, MT, P(qJ), P(q), P(k["JF"]))
    L.or_dgemm(b"T", b"N", Ms, MT * NF * 2, MG, 1.0, P(k["psi_w"]), MG, P(qJ), MG, 0.0, P(Qt), Ms)
    R.LmapToElement(Ns, NE, Ms, c.N_N, P(k["invmap"]), P(Q), P(Qt))
    R.LaddSFQ(Ns, NE, P(frk), P(S), P(F), P(Q))
    return frk


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_rhs_bitwise(c, order):
    R = oracle.ref_lib(*c)
    for seed, riem in ((1, False), (2, True)):
        mesh, t, prob, U = _setup(c, order, seed, riemann=riem)
        f_ref = _ref_rhs(R, prob, t, U)
        f_or = prob.rhs(U)
        assert np.array_equal(f_ref, f_or), f"max diff {np.abs(f_ref - f_or).max()}"
        assert np.abs(f_or).max() > 0


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
def test_solve_and_upa_bitwise(c):
    R = oracle.ref_lib(*c)
    mesh, t, prob, U = _setup(c, 2, 3)
    f = prob.rhs(U)
    du_ref = np.zeros_like(f)
    R.Lsolver(prob.cfg.N_s, prob.cfg.N_E, ctypes.c_double(0.0123), P(prob._keep["Minv"]), P(f), P(du_ref))
    assert np.array_equal(du_ref, prob.solve(0.0123, f))
    upa = np.zeros(prob.cfg.N_E * prob.cfg.N_s)
    R.LfindUPA(prob.cfg.N_s, prob.cfg.N_E, P(U), P(upa))
    assert prob.dt_from_cfl(U, 0.7) == 0.7 / upa[np.argmax(np.abs(upa))]


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2])
def test_hrl1D_bitwise(dim, order):
    c = (dim, MULTIFLUID, 0)
    R = oracle.ref_lib(*c)
    mesh, t, prob, U = _setup(c, order, 4, riemann=True)
    rng = np.random.default_rng(5)
    A = rng.standard_normal(U.size)
    Ns, NE, NF, NN = prob.cfg.N_s, prob.cfg.N_E, prob.nf, prob.cfg.N_N
    nb = prob._keep["neighbors"]
    if dim == 1:
        args = [(Ns, 1, 0)]
    else:
        args = [(order + 1, order + 1, 0), (order + 1, order + 1, 2)]
    for ns, sl, off in args:
        a1 = np.zeros_like(A)
        a2 = np.zeros_like(A)
        R.Lhrl1D(ns, NE, NF, NN, sl, P(nb), off, P(A), P(a1))
        oracle.lib().or_hrl1D(ns, NE, NF, NN, sl, P(nb), off, P(A), P(a2))
        assert np.array_equal(a1, a2)
        assert np.abs(a1).max() > 0


@pytest.mark.parametrize("c", [x for x in CONFIGS if x[1] == MULTIFLUID] + [x for x in CONFIGS_PS if x[2] == 1],
                         ids=cfg_id)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_sensors_and_hri_bitwise(c, order):
    R = oracle.ref_lib(*c)
    d = c[0]
    for seed in range(4):
        mesh, t, prob, U = _setup(c, order, 10 + seed, riemann=True)
        th = (0.02, 0.02, 0.02)[: 1 + seed % 3]
        sens_or, uavg_or = prob.sensing(U, th)
        # reference: blas mean + Lcalc_sensors (sensor.cc:11-36)
        uavg = np.zeros(prob.cfg.N_E * prob.nf)
        ones = np.ones(prob.cfg.N_s)
        oracle.lib().or_dgemm(b"N", b"N", 1, prob.cfg.N_E * prob.nf, prob.cfg.N_s, 1.0 / prob.cfg.N_s, P(ones), 1, P(U),
                              prob.cfg.N_s, 0.0, P(uavg), 1)
        assert np.array_equal(uavg, uavg_or)
        sens = np.zeros(prob.cfg.N_E, np.int32)
        on = [k < len(th) for k in range(3)]
        tv = list(th) + [0.0] * (3 - len(th))
        R.Lcalc_sensors(prob.cfg.N_E, prob.cfg.N_N, ctypes.c_bool(on[0]), ctypes.c_double(tv[0]), ctypes.c_bool(on[1]),
                        ctypes.c_double(tv[1]), ctypes.c_bool(on[2]), ctypes.c_double(tv[2]), P(prob._keep["neighbors"]),
                        P(uavg), P(sens))
        assert np.array_equal(sens, sens_or)
        # HRIlimiting (limiting.cc:596-634): hri1D on the flagged elements, then copy_detected, per direction
        U_ref = ref_hri_sweeps(R, prob, U, sens)
        U_or, s2 = prob.limit_hri(U, th)
        assert np.array_equal(s2, sens)
        assert np.array_equal(U_or, U_ref)


@pytest.mark.parametrize("c", CONFIGS + [x for x in CONFIGS_PS if x[1] == STIFFENED], ids=cfg_id)
def test_printer_formatter_bitwise(c):
    """PRINTER formatter (Lformater, printer_*.cu) both ways: conserved ->
    printed fields, and the restart direction."""
    R = oracle.ref_lib(*c)
    R.Lformater.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_bool]
    R.Lformater.restype = None
    for seed, riem in ((5, False), (6, True)):
        mesh, t, prob, U = _setup(c, 2, seed, riemann=riem)
        out_ref = np.zeros_like(U)
        Uc = U.copy()
        R.Lformater(prob.cfg.N_s, prob.cfg.N_E, P(Uc), P(out_ref), False)
        out = prob.format(U)
        assert np.array_equal(out, out_ref)
        back_ref = np.zeros_like(U)
        R.Lformater(prob.cfg.N_s, prob.cfg.N_E, P(back_ref), P(out_ref), True)
        assert np.array_equal(prob.unformat(out), back_ref)
        assert np.allclose(back_ref, U, rtol=1e-13, atol=1e-13)


BCS_LIM = {1: [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC)],
           2: [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE)]}


def _ref_m2(R, prob, t, U):
    """Limiting::M2limiting (limiting.cc:123-594, GAMNCON) composed from the
    reference launchers and netlib dgemm, in the reference's call order."""
    c, L = prob.cfg, oracle.lib()
    D, Ns, NE, NF, NN, n1 = c.D, c.N_s, c.N_E, prob.nf, c.N_N, c.N_s1D
    k = prob._keep
    U = U.copy()
    z = lambda: np.zeros(Ns * NE)
    rho, rhou, rhov, E, gam, pres, K, mono, lim, gLim, pLim, KLim, EMono, rhoeLim, ELim = (z() for _ in range(15))
    gemm = lambda M, X, Y: L.or_dgemm(b"N", b"N", Ns, NE, Ns, 1.0, P(M), Ns, P(X), Ns, 0.0, P(Y), Ns)
    get = lambda f, x: R.Lstridedcopy(NE, Ns, Ns * NF, Ns, f * Ns, 0, P(U), P(x))
    put = lambda f, x: R.Lstridedcopy(NE, Ns, Ns, Ns * NF, 0, f * Ns, P(x), P(U))
    if D == 1:
        def hr(x, Lm, back):
            gemm(k["Lag2Mono"], x, mono)
            R.Lhrl1D(Ns, NE, 1, NN, 1, P(k["neighbors"]), 0, P(mono), P(Lm))
            if back:
                gemm(k["Mono2Lag"], Lm, x)
        get(0, rho); hr(rho, lim, True)
        get(1, rhou); hr(rhou, lim, True)
        get(2, E); gemm(k["Lag2Mono"], E, EMono)
        get(3, gam); hr(gam, gLim, True)
        R.Lpressure(Ns, NE, P(U), P(pres)); hr(pres, pLim, False)
        R.Lkinetic_energy1D(Ns, NE, P(rho), P(rhou), P(K)); gemm(k["Lag2Mono"], K, KLim)
        R.Linternal_energy_multifluid(Ns, NE, 1, P(pLim), P(gLim), P(rhoeLim))
        R.Lreconstruct_energy(Ns, NE, 1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
        gemm(k["Mono2Lag"], ELim, E)
        for f, x in enumerate((rho, rhou, E, gam)):
            put(f, x)
        return U

    def hr(L2M, M2L, off, x, Lm, back):
        gemm(k[L2M], x, mono)
        R.Lhrl1D(n1, NE, 1, NN, n1, P(k["neighbors"]), off, P(mono), P(Lm))
        if back:
            gemm(k[M2L], Lm, x)
    for f, x in ((0, rho), (1, rhou), (2, rhov)):
        get(f, x); hr("Lag2MonoX", "MonoX2Lag", 0, x, lim, True)
    get(4, gam); hr("Lag2MonoX", "MonoX2Lag", 0, gam, gLim, True)
    R.Lpressure(Ns, NE, P(U), P(pres)); hr("Lag2MonoX", "MonoX2Lag", 0, pres, pLim, True)
    R.Lkinetic_energy2D(Ns, NE, P(rho), P(rhou), P(rhov), P(K)); gemm(k["Lag2MonoX"], K, KLim)
    R.Linternal_energy_multifluid(n1, NE, n1, P(pLim), P(gLim), P(rhoeLim))
    get(3, E); gemm(k["Lag2MonoX"], E, EMono)
    R.Lreconstruct_energy(n1, NE, n1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
    gemm(k["MonoX2Lag"], ELim, E)
    for x in (rho, rhou, rhov):
        hr("Lag2MonoY", "MonoY2Lag", 2, x, lim, True)
    hr("Lag2MonoY", "MonoY2Lag", 2, gam, gLim, True)
    hr("Lag2MonoY", "MonoY2Lag", 2, pres, pLim, False)
    R.Lkinetic_energy2D(Ns, NE, P(rho), P(rhou), P(rhov), P(K)); gemm(k["Lag2MonoY"], K, KLim)
    R.Linternal_energy_multifluid(n1, NE, n1, P(pLim), P(gLim), P(rhoeLim))
    gemm(k["Lag2MonoY"], E, EMono)
    R.Lreconstruct_energy(n1, NE, n1, P(rhoeLim), P(KLim), P(EMono), P(ELim))
    gemm(k["MonoY2Lag"], ELim, E)
    for f, x in enumerate((rho, rhou, rhov, E, gam)):
        put(f, x)
    return U


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_m2_and_pr_limiting_bitwise(dim, order):
    """M2limiting, PRlimiting and PRIlimiting (two-fluid) against the reference launchers."""
    c = (dim, MULTIFLUID, 1)
    R = oracle.ref_lib(*c)
    for k, bc in enumerate(BCS_LIM[dim]):
        mesh, t, prob, U = _setup(c, order, 30 + k, riemann=True, bc=bc)
        x = np.arange(U.size, dtype=np.float64)
        U = U * (1.0 + 0.01 * np.sin(0.37 * x + k))
        U_ref = _ref_m2(R, prob, t, U)
        assert np.array_equal(prob.limit_m2(U), U_ref)
        assert not np.array_equal(U_ref, U)
        # PRlimiting (limiting.cc:102-121): Cons2Prim, HRlimiting composed from the reference launchers, Prim2Cons
        V = U.copy()
        R.LCons2Prim(prob.cfg.N_s, prob.cfg.N_E, P(V))
        V = ref_hr(R, prob, V)
        R.LPrim2Cons(prob.cfg.N_s, prob.cfg.N_E, P(V))
        assert np.array_equal(prob.limit_pr(U), V)
        # PRIlimiting (limiting.cc:636-662): sensing (pinned above), Cons2Prim, hri1D + copy_detected, Prim2Cons
        th = (0.01, 0.01, 0.01)
        sens, _ = prob.sensing(U, th)
        assert sens.any()
        cfg, kk = prob.cfg, prob._keep
        V = U.copy()
        R.LCons2Prim(cfg.N_s, cfg.N_E, P(V))
        sweeps = ([("Lag2Mono", "Mono2Lag", cfg.N_s, 1, 0)] if dim == 1 else
                  [("Lag2MonoX", "MonoX2Lag", cfg.N_s1D, cfg.N_s1D, 0), ("Lag2MonoY", "MonoY2Lag", cfg.N_s1D, cfg.N_s1D, 2)])
        for l2m, m2l, n1, sl, off in sweeps:
            Vn = np.zeros_like(V)
            R.Lhri1D(cfg.N_s, cfg.N_E, cfg.N_N, P(kk["neighbors"]), n1, sl, off, P(kk[l2m]), P(kk[m2l]), P(sens), P(V),
                     P(Vn))
            R.Lcopy_detected(cfg.N_s, cfg.N_E, P(sens), P(Vn), P(V))
        R.LPrim2Cons(cfg.N_s, cfg.N_E, P(V))
        W, s2 = prob.limit_pri(U, th)
        assert np.array_equal(s2, sens) and np.array_equal(W, V)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_m2i_limiting_bitwise(dim, order):
    """M2Ilimiting (two-fluid): sensing, then Lm2i1D + Lcopy_detected per direction, against the oracle."""
    c = (dim, MULTIFLUID, 1)
    R = oracle.ref_lib(*c)
    for k, bc in enumerate(BCS_LIM[dim]):
        mesh, t, prob, U = _setup(c, order, 50 + k, riemann=True, bc=bc)
        x = np.arange(U.size, dtype=np.float64)
        U = U * (1.0 + 0.01 * np.sin(0.41 * x + k))
        th = (0.01, 0.01, 0.01)
        sens, _ = prob.sensing(U, th)
        assert sens.any()
        cfg, kk = prob.cfg, prob._keep
        V = U.copy()
        sweeps = ([("Lag2Mono", "Mono2Lag", cfg.N_s, 1, 0)] if dim == 1 else
                  [("Lag2MonoX", "MonoX2Lag", cfg.N_s1D, cfg.N_s1D, 0), ("Lag2MonoY", "MonoY2Lag", cfg.N_s1D, cfg.N_s1D, 2)])
        for l2m, m2l, n1, sl, off in sweeps:
            Vn = np.zeros_like(V)
            R.Lm2i1D(cfg.N_s, cfg.N_E, cfg.N_N, P(kk["neighbors"]), n1, sl, off, P(kk[l2m]), P(kk[m2l]), P(sens), P(V),
                     P(Vn))
            # one sweep of the oracle on the same input
            Wn = np.zeros_like(V)
            oracle.lib().or_m2i1D(ctypes.byref(cfg), P(kk["neighbors"]), n1, sl, off, P(kk[l2m]), P(kk[m2l]), P(sens),
                                  P(V), P(Wn))
            assert np.array_equal(Wn, Vn)
            R.Lcopy_detected(cfg.N_s, cfg.N_E, P(sens), P(Vn), P(V))
        W, s2 = prob.limit_m2i(U, th)
        assert np.array_equal(s2, sens) and np.array_equal(W, V)
        assert not np.array_equal(W, U)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_p0i_limiting_bitwise(dim, order):
    """P0Ilimiting: sensing, Lp0i + Lcopy_detected against the oracle.  The
    reference sets refArea = 4 for quadrilaterals (main.cc:286) and leaves it
    unset for segments; the launcher takes it as an argument, so 1D is pinned
    with the segment's length 2 that the oracle uses."""
    c = (dim, MULTIFLUID, 1)
    R = oracle.ref_lib(*c)
    R.Lp0i.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp]
    for k, bc in enumerate(BCS_LIM[dim]):
        mesh, t, prob, U = _setup(c, order, 60 + k, riemann=True, bc=bc)
        cfg, kk = prob.cfg, prob._keep
        th = (0.01, 0.01, 0.01)
        sens, _ = prob.sensing(U, th)
        assert sens.any() and not sens.all()
        area = oracle.lib().or_ref_area(ctypes.byref(cfg))
        assert area == (4.0 if dim == 2 else 2.0)
        Vn = np.zeros_like(U)
        R.Lp0i(cfg.N_s, cfg.N_E, cfg.N_G, area, P(sens), P(kk["phi_w"]), P(U), P(Vn))
        V = U.copy()
        R.Lcopy_detected(cfg.N_s, cfg.N_E, P(sens), P(Vn), P(V))
        W, s2 = prob.limit_p0i(U, th)
        assert np.array_equal(s2, sens) and np.array_equal(W, V)
        # each flagged element's fields are constant; the others are untouched
        W3 = W.reshape(cfg.N_E, prob.nf, cfg.N_s)
        f = sens != 0
        assert np.array_equal(W3[f], np.repeat(W3[f][:, :, :1], cfg.N_s, axis=2))
        assert np.array_equal(W3[~f], U.reshape(cfg.N_E, prob.nf, cfg.N_s)[~f])

