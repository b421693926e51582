"""Closed-form checks of the oracle's limiters that hold without the
reference's sources (the reference-launcher pins are in
test_oracle_vs_reference.py)."""
import ctypes

import numpy as np
import pytest

import oracle
from dg_amd import MULTIFLUID, problems
from tests.helpers import small_mesh


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _setup(c, order, seed):
    d, ph, fl = c
    mesh = small_mesh(d, order, ph)
    t = mesh.tables()
    return mesh, t, oracle.Problem(t, ph, fl), problems.riemann_state(mesh, ph, seed)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_p0i_average_is_the_cell_mean(dim, order):
    """For a state that is a polynomial of degree <= order in each element, the
    p0i value equals the exact cell mean: here the mean of x (and y) over each
    element is its centre, so the limited field is the element centre's value."""
    c = (dim, MULTIFLUID, 1)
    mesh, t, prob, U = _setup(c, order, 70)
    cfg = prob.cfg
    xyz = np.asarray(t["xyz"], dtype=np.float64).reshape(cfg.N_E, cfg.N_s, dim)
    U3 = np.ones((cfg.N_E, prob.nf, cfg.N_s))
    lin = 1.0 + 0.25 * xyz[:, :, 0] + (0.125 * xyz[:, :, 1] if dim == 2 else 0.0)
    U3[:, 1, :] = lin
    flags = np.ones(cfg.N_E, np.int32)
    Unew = np.zeros(U3.size)
    oracle.lib().or_p0i(ctypes.byref(cfg), P(prob._keep["phi_w"]), oracle.lib().or_ref_area(ctypes.byref(cfg)),
                        P(flags), P(np.ascontiguousarray(U3.ravel())), P(Unew))
    got = Unew.reshape(cfg.N_E, prob.nf, cfg.N_s)[:, 1, 0]
    cen = xyz.mean(axis=1)  # the nodal mean of x is the centre for the symmetric node sets
    want = 1.0 + 0.25 * cen[:, 0] + (0.125 * cen[:, 1] if dim == 2 else 0.0)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    assert np.allclose(Unew.reshape(cfg.N_E, prob.nf, cfg.N_s)[:, 0, :], 1.0, rtol=1e-14)
