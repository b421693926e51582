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
    }
    for n, a in sig.items():
        getattr(lib, n).argtypes = a
    return lib


def D(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def P(t):
    return vp(t.data_ptr())


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
    if dim == 1:
        U_ref, s2 = prob.limit_hri(U, th)
        assert np.array_equal(s2, sens_ref)
        assert np.array_equal(H(Ud), U_ref)
    assert (sens_ref != 0).any()
