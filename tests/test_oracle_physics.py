"""Closed-form checks of the oracle's passive-scalar formatter, whose
reference source (printer_{oned,twod}_passive.cu) does not compile on its own
(it names constants::GLOBAL_GAMMA without including constants.h): states are
built from chosen primitive values, and the printed fields must give those
values back, and the restart direction the conserved state."""
import numpy as np
import pytest

import oracle
from dg_amd import PASSIVE, ROE
from tests.helpers import small_mesh


@pytest.mark.parametrize("dim", [1, 2])
def test_passive_formatter_closed_form(dim):
    mesh = small_mesh(dim, 2, PASSIVE)
    prob = oracle.Problem(mesh.tables(), PASSIVE, ROE)
    NE, NF, Ns = prob.cfg.N_E, prob.nf, prob.cfg.N_s
    rng = np.random.default_rng(3)
    rho, p = rng.uniform(0.5, 2.0, (2, NE, Ns))
    vel = rng.uniform(-1.0, 1.0, (dim, NE, Ns))
    phic, phinc = rng.uniform(0.0, 1.0, (2, NE, Ns))
    ke = 0.5 * rho * (vel ** 2).sum(axis=0)
    U = np.zeros((NE, NF, Ns))
    U[:, 0] = rho
    U[:, 1:1 + dim] = (rho * vel).transpose(1, 0, 2)
    U[:, dim + 1] = p / 0.4 + ke  # gamma = 1.4
    U[:, dim + 2] = rho * phic
    U[:, dim + 3] = phinc
    out = prob.format(U.ravel()).reshape(NE, NF, Ns)
    want = [rho] + list(vel) + [p, phic, phinc]
    for f, w in enumerate(want):
        assert np.allclose(out[:, f], w, rtol=1e-12, atol=1e-12), f
    back = prob.unformat(out.ravel()).reshape(NE, NF, Ns)
    assert np.allclose(back, U, rtol=1e-12, atol=1e-12)
