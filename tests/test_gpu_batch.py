"""GPU parity of the batched ensemble (dgx_batched_rk4) against the oracle.

Every problem of a small uneven ensemble (orders 1-3, different element
counts, two-fluid Riemann states, Roe flux, HR limiter) is run by the oracle
on its own, in CFL mode with the initial limiting and a step cap, and must
match the batched run to 1e-11 relative L2/Linf per field, with the same
per-problem time.  The graph is replayed and the autotuner's choice must
not change results beyond that tolerance.
"""
import numpy as np
import pytest
import torch

import oracle
from dg_amd import FARFIELD, LIMIT_HR, MULTIFLUID, REFLECTIVE, ROE, Batch, Config, Mesh, problems
from tests.helpers import per_field_rel

pytestmark = pytest.mark.gpu


def _ensemble(n, seed, ne_range=(20, 90), bc=(FARFIELD, FARFIELD)):
    ne, p = problems.ensemble_spec(n, seed=seed, ne_range=ne_range)
    meshes = [Mesh(1, p[k], ne[k], 1, -1.0, 1.0, bc=bc, nfields=4) for k in range(n)]
    # state seeds 900+k: every problem of these ensembles survives the run in
    # the oracle (strong random Riemann problems can lose positivity)
    states = [problems.two_fluid_riemann_1d(meshes[k], 900 + k) for k in range(n)]
    return ne, p, meshes, states


def _oracle_run(meshes, states, p, cfl, nsteps):
    out = []
    for k, m in enumerate(meshes):
        prob = oracle.Problem(m.tables(), MULTIFLUID, ROE)
        U, n, T, _ = prob.rk4(states[k], 1, cfl * m.dx / (2 * p[k] + 1), nsteps=nsteps, tf=1e30, limiter=LIMIT_HR)
        out.append((U, n, T, prob))
    return out


@pytest.mark.parametrize("bc", [(FARFIELD, FARFIELD), (REFLECTIVE, FARFIELD)])
def test_batched_rk4_matches_oracle(device, bc):
    n, cfl, nsteps = 12, 0.4, 20
    ne, p, meshes, states = _ensemble(n, seed=11, bc=bc)
    ref = _oracle_run(meshes, states, p, cfl, nsteps)
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p, bc=bc)
    U = torch.from_numpy(batch.pack(states)).to(device)
    T = torch.zeros(n, dtype=torch.float64, device=device)
    status = torch.zeros(n, dtype=torch.int32, device=device)
    batch.rk4(U, cfl, nsteps, T, status)
    torch.cuda.synchronize()
    Uh = batch.unpack(U.cpu().numpy())
    Th, sh = T.cpu().numpy(), status.cpu().numpy()
    for k in range(n):
        U_ref, n_ref, T_ref, prob = ref[k]
        assert n_ref == nsteps and sh[k] == 0, k
        assert abs(Th[k] - T_ref) <= 1e-12 * T_ref, (k, Th[k], T_ref)
        for l2, li in per_field_rel(Uh[k], U_ref, prob.nf, prob.cfg.N_s):
            assert l2 < 1e-11 and li < 1e-11, (k, p[k], ne[k], l2, li)


def test_batched_graph_replay_and_autotune(device):
    """Two replays of the captured graph equal one run of twice the steps;
    workgroups of 64, 128, 256, 512 and 1024 threads per problem give the
    same bits, and the autotuner's choice agrees."""
    n, cfl = 9, 0.4
    ne, p, meshes, states = _ensemble(n, seed=5)
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p)
    U0 = torch.from_numpy(batch.pack(states)).to(device)

    def run(sizes, chunks, steps):
        if sizes:
            batch.set_block_sizes(sizes)
        U = U0.clone()
        T = torch.zeros(n, dtype=torch.float64, device=device)
        st = torch.zeros(n, dtype=torch.int32, device=device)
        for c in range(chunks):
            batch.rk4(U, cfl, steps, T, st, limit_initial=(c == 0))
        torch.cuda.synchronize()
        return U.cpu().numpy(), T.cpu().numpy(), st.cpu().numpy()

    Ua, Ta, sa = run([64], 1, 16)  # one workgroup of 64 threads per problem
    assert batch.block_sizes[:2] == [64, n]
    Ub, Tb, sb = run([64], 2, 8)
    assert np.array_equal(Ua, Ub) and np.array_equal(Ta, Tb) and np.array_equal(sa, sb)
    Us, Ts, ss = run([1024], 2, 8)  # the same arithmetic per element, whatever the workgroup size
    assert np.array_equal(sa, ss) and np.array_equal(Ta, Ts) and np.array_equal(Ua, Us)
    for bs in (128, 256, 512):  # the three-wave instance (<= 256 threads) and the two-wave one
        Uw, Tw, sw = run([bs], 1, 16)
        assert np.array_equal(sa, sw) and np.array_equal(Ta, Tw) and np.array_equal(Ua, Uw), bs
    batch.autotune(U0, cfl)
    sizes = batch.block_sizes
    assert sizes[0] in (64, 128, 256, 512, 1024) and sizes[1] == n
    Uc, Tc, sc = run(None, 1, 16)
    assert np.array_equal(sa, sc)
    scale = np.abs(Ua).max()
    assert np.abs(Uc - Ua).max() <= 1e-12 * scale
    assert np.abs(Tc - Ta).max() <= 1e-13 * np.abs(Ta).max()


def test_batched_failure_stops_one_problem(device):
    """A problem with a negative initial pressure gets a NaN time step: it
    stops with status 1 and leaves the other problems untouched."""
    n, cfl, nsteps = 6, 0.4, 10
    ne, p, meshes, states = _ensemble(n, seed=11)
    bad = 2
    states[bad] = states[bad].copy()
    Ns = p[bad] + 1
    states[bad].reshape(ne[bad], 4, Ns)[:, 2, :] *= -1.0  # E < 0: negative pressure
    ref = _oracle_run(meshes, states, p, cfl, nsteps)
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p)
    U = torch.from_numpy(batch.pack(states)).to(device)
    T = torch.zeros(n, dtype=torch.float64, device=device)
    status = torch.zeros(n, dtype=torch.int32, device=device)
    batch.rk4(U, cfl, nsteps, T, status)
    torch.cuda.synchronize()
    sh = status.cpu().numpy()
    Uh = batch.unpack(U.cpu().numpy())
    assert sh[bad] == 1 and ref[bad][1] == -1
    for k in range(n):
        if k == bad:
            continue
        U_ref, n_ref, T_ref, prob = ref[k]
        assert sh[k] == 0 and n_ref == nsteps
        for l2, li in per_field_rel(Uh[k], U_ref, prob.nf, prob.cfg.N_s):
            assert l2 < 1e-11 and li < 1e-11, (k, l2, li)


def test_large_workgroups_match_oracle(device):
    """Problems of 150-400 elements (more elements than lanes at 256
    threads, fewer at 1024): every problem equals its own oracle run to
    1e-11, and workgroups of 256 and 1024 threads give the same bits."""
    n, cfl, nsteps = 4, 0.4, 12
    ne, p, meshes, states = _ensemble(n, seed=7, ne_range=(150, 400), bc=(REFLECTIVE, FARFIELD))
    ref = _oracle_run(meshes, states, p, cfl, nsteps)
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p, bc=(REFLECTIVE, FARFIELD))
    U0 = torch.from_numpy(batch.pack(states)).to(device)
    out = []
    for block in (256, 1024):
        batch.set_block_sizes([block])
        U = U0.clone()
        T = torch.zeros(n, dtype=torch.float64, device=device)
        status = torch.zeros(n, dtype=torch.int32, device=device)
        batch.rk4(U, cfl, nsteps, T, status)
        torch.cuda.synchronize()
        out.append((U.cpu().numpy(), T.cpu().numpy(), status.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    Uh = batch.unpack(out[1][0])
    for k in range(n):
        U_ref, n_ref, T_ref, prob = ref[k]
        assert n_ref == nsteps and out[1][2][k] == 0, k
        assert abs(out[1][1][k] - T_ref) <= 1e-12 * T_ref
        for l2, li in per_field_rel(Uh[k], U_ref, prob.nf, prob.cfg.N_s):
            assert l2 < 1e-11 and li < 1e-11, (k, p[k], ne[k], l2, li)


@pytest.mark.parametrize("limiter,limit_initial", [(LIMIT_HR, True), (LIMIT_HR, False), (0, False)])
def test_cut_problems_equal_uncut(device, limiter, limit_initial):
    """Problems cut into chunks on several workgroups (exchanging their end
    elements every stage and gathering Dt every step) give the bits of the
    same problems on one workgroup each: state, time and status, over two
    graph launches; with the limiter, every problem also equals its own
    oracle run to 1e-11."""
    n, cfl, nsteps = 6, 0.4, 10
    ne, p, meshes, states = _ensemble(n, seed=23, ne_range=(120, 400), bc=(REFLECTIVE, FARFIELD))
    if not limiter:  # smooth states: the Riemann states lose positivity without the limiter
        states = [problems.smooth_random_state(m, MULTIFLUID, seed=40 + k) for k, m in enumerate(meshes)]
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p, bc=(REFLECTIVE, FARFIELD))
    U0 = torch.from_numpy(batch.pack(states)).to(device)
    out = {}
    for min_chunk in (0, 16):
        batch.set_min_chunk(min_chunk)
        batch.set_block_sizes([256])
        sizes = batch.block_sizes
        if min_chunk == 0:
            assert sizes[1:] == [n, 0]
        else:
            assert sizes[2] > 0 and sizes[1] > n, sizes
        U = U0.clone()
        T = torch.zeros(n, dtype=torch.float64, device=device)
        st = torch.zeros(n, dtype=torch.int32, device=device)
        for c in range(2):
            batch.rk4(U, cfl, nsteps // 2, T, st, limiter=limiter, limit_initial=limit_initial and c == 0)
        torch.cuda.synchronize()
        out[min_chunk] = (U.cpu().numpy(), T.cpu().numpy(), st.cpu().numpy())
    (Ua, Ta, sa), (Ub, Tb, sb) = out[0], out[16]
    assert not sa.any() and not sb.any()
    assert np.array_equal(Ta, Tb) and np.array_equal(Ua, Ub)
    if limiter == LIMIT_HR and limit_initial:
        ref = _oracle_run(meshes, states, p, cfl, nsteps)
        Uh = batch.unpack(Ub)
        for k in range(n):
            U_ref, n_ref, T_ref, prob = ref[k]
            assert n_ref == nsteps and abs(Tb[k] - T_ref) <= 1e-12 * T_ref
            for l2, li in per_field_rel(Uh[k], U_ref, prob.nf, prob.cfg.N_s):
                assert l2 < 1e-11 and li < 1e-11, (k, p[k], ne[k], l2, li)
