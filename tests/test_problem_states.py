"""Host-side checks of the workload generators and of the state-size guard of
the Python binding (CPU only): every generator returns one field per field of
the mesh for every physics, and a state of the wrong size, dtype or layout is
refused before anything reaches the library."""
import numpy as np
import pytest
import torch

from dg_amd import MULTIFLUID, PASSIVE, SINGLEFLUID, STIFFENED, _state_ptr, nfields, problems

PHYSICS = [SINGLEFLUID, MULTIFLUID, PASSIVE, STIFFENED]


@pytest.mark.parametrize("ph", PHYSICS)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_sod_state_has_every_field(ph, order):
    mesh = problems.sod_mesh(40, order, ph)
    s = mesh.sizes
    U = problems.sod_state(mesh, ph)
    assert U.size == s["N_E"] * nfields(1, ph) * s["N_s"]
    assert np.isfinite(U).all()
    F = U.reshape(s["N_E"], nfields(1, ph), s["N_s"])
    assert (F[:, 0] > 0).all()  # density
    if ph == STIFFENED:
        assert np.array_equal(F[:, 4], np.zeros_like(F[:, 4]))  # pinf = 0: beta = 0
    if ph == PASSIVE:
        assert set(np.unique(F[:, 4])) == {0.0, 1.0}


def test_sod_state_keeps_its_single_and_two_fluid_values():
    """The single- and two-fluid Sod states are the reference's (rho, rho u,
    E[, 1/(gamma-1)]) with gamma = 1.4, split at x = 0."""
    for ph in (SINGLEFLUID, MULTIFLUID):
        mesh = problems.sod_mesh(40, 2, ph)
        F = problems.sod_state(mesh, ph).reshape(40, nfields(1, ph), -1)
        assert np.array_equal(F[0, 0], np.ones(3)) and np.array_equal(F[-1, 0], np.full(3, 0.125))
        assert np.array_equal(F[0, 2], np.full(3, 1.0 / (1.4 - 1.0)))
        assert np.array_equal(F[-1, 2], np.full(3, 0.1 / (1.4 - 1.0)))
        if ph == MULTIFLUID:
            assert np.array_equal(F[:, 3], np.full((40, 3), 1.0 / (1.4 - 1.0)))


@pytest.mark.parametrize("ph", PHYSICS)
def test_riemann_and_smooth_states_have_every_field(ph):
    for d, mesh in ((1, problems.sod_mesh(16, 2, ph)),
                    (2, problems.Mesh(2, 2, 8, 4, 0.0, 1.0, 0.0, 1.0, nfields=nfields(2, ph)))):
        n = mesh.sizes["N_E"] * nfields(d, ph) * mesh.sizes["N_s"]
        assert problems.riemann_state(mesh, ph, seed=1).size == n
        assert problems.smooth_random_state(mesh, ph, seed=1).size == n


def test_pack_refuses_a_wrong_field_count():
    mesh = problems.sod_mesh(8, 1, PASSIVE)
    x, _ = problems.node_xy(mesh)
    with pytest.raises(ValueError):
        problems._pack(mesh, [x, x, x])


def test_state_guard():
    n = 24
    assert _state_ptr(torch.zeros(n, dtype=torch.float64), n, "t") != 0
    for bad in (torch.zeros(n + 1, dtype=torch.float64), torch.zeros(n, dtype=torch.float32),
                torch.zeros(2 * n, dtype=torch.float64)[::2]):
        with pytest.raises(ValueError):
            _state_ptr(bad, n, "t")
    assert _state_ptr(None, n, "t") is None
    assert _state_ptr(1234, n, "t") == 1234
