"""Pin the oracle's restatements of the reference's helper launchers to the
reference's own kernel files (compiled as CPU loops into oracle/_ref), bit
for bit: Lstridedcopy, Lreconstruct_energy, Linternal_energy_multifluid /
_stiffened (limiting_kernels.cu:51-243), Lpressure, Lkinetic_energy1D / 2D
(physics.cu:814-960), Laverage_cell_p0 (rk_kernels.cu:61-92), LChangeBasis,
LCons2Prim, LPrim2Cons (limiting_kernels.cu:918-1068) and Lformat_sensor
(printer_sensor.cu:19-43).  The GPU launchers of the same names are checked
against these restatements in tests/test_gpu_dropin.py."""
import ctypes

import numpy as np
import pytest

import oracle
from dg_amd import MULTIFLUID, PASSIVE, SINGLEFLUID, STIFFENED, problems
from tests.helpers import CONFIGS, CONFIGS_PS, cfg_id, small_mesh

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="reference sources not present: _ref not built")
vp = ctypes.c_void_p


def P(a):
    return a.ctypes.data_as(vp)


def _prob(c, order=2, seed=3):
    d, ph, fl = c
    mesh = small_mesh(d, order, ph)
    prob = oracle.Problem(mesh.tables(), ph, fl)
    return mesh, prob, problems.riemann_state(mesh, ph, seed)


def test_stridedcopy():
    R, L = oracle.ref_lib(1, MULTIFLUID, 0), oracle.lib()
    rng = np.random.default_rng(1)
    A = rng.standard_normal(400)
    for args in ((2, 3, 9, 3, 0, 0), (7, 4, 20, 5, 3, 1), (12, 3, 12, 3, 6, 0), (5, 9, 45, 9, 9, 2)):
        B1 = np.full(200, 7.0)
        B2 = np.full(200, 7.0)
        R.Lstridedcopy(*args, P(A), P(B1))
        L.or_stridedcopy(*args, P(A), P(B2))
        assert np.array_equal(B1, B2) and not np.all(B1 == 7.0)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_energy_reconstruction_pieces(order):
    R, L = oracle.ref_lib(2, STIFFENED, 1), oracle.lib()
    rng = np.random.default_rng(order)
    n1 = order + 1
    NE = 23
    for Ns, sl in ((n1, 1), (n1, n1)):
        n = Ns * sl * NE
        p, g, b, K, E = (rng.standard_normal(n) for _ in range(5))
        r1, r2 = np.zeros(n), np.zeros(n)
        R.Linternal_energy_multifluid(Ns, NE, sl, P(p), P(g), P(r1))
        L.or_internal_energy(Ns, NE, sl, P(p), P(g), None, P(r2))
        assert np.array_equal(r1, r2)
        R.Linternal_energy_stiffened(Ns, NE, sl, P(p), P(g), P(b), P(r1))
        L.or_internal_energy(Ns, NE, sl, P(p), P(g), P(b), P(r2))
        assert np.array_equal(r1, r2)
        e1, e2 = np.zeros(n), np.zeros(n)
        R.Lreconstruct_energy(Ns, NE, sl, P(r1), P(K), P(E), P(e1))
        L.or_reconstruct_energy(Ns, NE, sl, P(r1), P(K), P(E), P(e2))
        assert np.array_equal(e1, e2)


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
def test_pressure_and_kinetic_energy(c):
    R, L = oracle.ref_lib(*c), oracle.lib()
    mesh, prob, U = _prob(c)
    NE, Ns, D = prob.cfg.N_E, prob.cfg.N_s, c[0]
    p1, p2 = np.zeros(NE * Ns), np.zeros(NE * Ns)
    R.Lpressure(Ns, NE, P(U), P(p1))
    L.or_pressure(ctypes.byref(prob.cfg), P(U), P(p2))
    assert np.array_equal(p1, p2) and np.all(p1 > 0)
    V = U.reshape(NE, prob.nf, Ns)
    rho, ru = np.ascontiguousarray(V[:, 0]).ravel(), np.ascontiguousarray(V[:, 1]).ravel()
    k1, k2 = np.zeros(NE * Ns), np.zeros(NE * Ns)
    if D == 1:
        R.Lkinetic_energy1D(Ns, NE, P(rho), P(ru), P(k1))
        L.or_kinetic_energy(Ns, NE, P(rho), P(ru), None, P(k2))
    else:
        rv = np.ascontiguousarray(V[:, 2]).ravel()
        R.Lkinetic_energy2D(Ns, NE, P(rho), P(ru), P(rv), P(k1))
        L.or_kinetic_energy(Ns, NE, P(rho), P(ru), P(rv), P(k2))
    assert np.array_equal(k1, k2)


@pytest.mark.parametrize("c", CONFIGS + CONFIGS_PS, ids=cfg_id)
def test_average_cell_p0_and_change_basis(c):
    R, L = oracle.ref_lib(*c), oracle.lib()
    mesh, prob, U = _prob(c, order=3)
    NE, Ns, NF = prob.cfg.N_E, prob.cfg.N_s, prob.nf
    rng = np.random.default_rng(7)
    DU1 = rng.standard_normal(U.size)
    DU2 = DU1.copy()
    R.Laverage_cell_p0(Ns, NE, P(DU1))
    L.or_average_cell_p0(ctypes.byref(prob.cfg), P(DU2))
    assert np.array_equal(DU1, DU2)
    for s1, s2 in ((Ns, Ns), (Ns - 1, Ns), (Ns, Ns - 2)):
        T = rng.standard_normal(NE * s1 * s2)
        X = rng.standard_normal(NE * NF * s2)
        o1, o2 = np.zeros(NE * NF * s1), np.zeros(NE * NF * s1)
        R.LChangeBasis(s1, s2, NE, P(T), P(X), P(o1))
        L.or_change_basis(ctypes.byref(prob.cfg), s1, s2, P(T), P(X), P(o2))
        assert np.array_equal(o1, o2)


@pytest.mark.parametrize("c", [x for x in CONFIGS + CONFIGS_PS if x[1] != SINGLEFLUID], ids=cfg_id)
def test_cons_prim(c):
    """Cons2Prim / Prim2Cons; single-fluid states have no alpha field (the
    reference reads past the element there) and are left out."""
    R, L = oracle.ref_lib(*c), oracle.lib()
    mesh, prob, U = _prob(c)
    NE, Ns = prob.cfg.N_E, prob.cfg.N_s
    V1, V2 = U.copy(), U.copy()
    R.LCons2Prim(Ns, NE, P(V1))
    L.or_cons_prim(ctypes.byref(prob.cfg), P(V2), 1)
    assert np.array_equal(V1, V2) and not np.array_equal(V1, U)
    R.LPrim2Cons(Ns, NE, P(V1))
    L.or_cons_prim(ctypes.byref(prob.cfg), P(V2), 0)
    assert np.array_equal(V1, V2)
    assert np.allclose(V1, U, rtol=1e-12, atol=1e-12)


def test_format_sensor():
    R, L = oracle.ref_lib(2, MULTIFLUID, 1), oracle.lib()
    sens = np.array([0, 1, 2, 3, 0, 2, 1], np.int32)
    o1, o2 = np.zeros(7 * 9), np.zeros(7 * 9)
    R.Lformat_sensor(9, 7, P(sens), P(o1))
    L.or_format_sensor(9, 7, P(sens), P(o2))
    assert np.array_equal(o1, o2) and np.array_equal(o1.reshape(7, 9)[:, 4], sens)
