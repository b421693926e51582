"""Closed-form checks of the oracle's DG_SOLVER::conservation restatement
(or_conservation, dg_solver.h:367-389).

The reference's conservation() is host code inside the DG_SOLVER class and
does not build from its own sources here, so it is pinned by what the
mathematics fixes: for a polynomial field of degree <= p in each coordinate,
the nodal interpolant is the field itself and the volume rule (order 2p+1)
integrates it exactly, so each total is the field's exact integral over the
domain; a field shifted by a constant c adds c times the area; and the
running sum follows the reference's order (elements, then Gauss points).
"""
import math

import numpy as np
import pytest

import oracle
from dg_amd import FARFIELD, MULTIFLUID, PERIODIC, REFLECTIVE, ROE, Mesh, problems


def _mesh(dim, order, nx=9, ny=5):
    if dim == 1:
        return Mesh(1, order, nx, 1, -1.0, 2.0, bc=(FARFIELD, REFLECTIVE), nfields=4)
    return Mesh(2, order, nx, ny, 0.0, 1.2, -0.3, 0.7, bc=(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=5)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2])
def test_conservation_integrates_polynomials_exactly(dim, order):
    mesh = _mesh(dim, order)
    prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
    x, y = problems.node_xy(mesh)
    nf = prob.nf
    rng = np.random.default_rng(10 * dim + order)
    fields, exact = [], []
    for f in range(nf):
        if dim == 1:
            a = rng.uniform(-1, 1, order + 1)
            fields.append(sum(a[k] * x ** k for k in range(order + 1)))
            # integral of x^k over [-1, 2]
            exact.append(sum(a[k] * (2.0 ** (k + 1) - (-1.0) ** (k + 1)) / (k + 1) for k in range(order + 1)))
        else:
            a = rng.uniform(-1, 1, (order + 1, order + 1))
            fields.append(sum(a[i, j] * x ** i * y ** j for i in range(order + 1) for j in range(order + 1)))
            ix = [(1.2 ** (i + 1) - 0.0) / (i + 1) for i in range(order + 1)]
            iy = [(0.7 ** (j + 1) - (-0.3) ** (j + 1)) / (j + 1) for j in range(order + 1)]
            exact.append(sum(a[i, j] * ix[i] * iy[j] for i in range(order + 1) for j in range(order + 1)))
    U = np.ascontiguousarray(np.stack(fields, axis=1).reshape(-1))
    tot = prob.conservation(U)
    for f in range(nf):
        assert abs(tot[f] - exact[f]) <= 1e-13 * max(1.0, abs(exact[f])), (f, tot[f], exact[f])
    # a constant shift adds the constant times the measure of the domain
    area = 3.0 if dim == 1 else 1.2 * 1.0
    tot2 = prob.conservation(U + 2.5)
    for f in range(nf):
        assert abs(tot2[f] - tot[f] - 2.5 * area) <= 1e-13 * max(1.0, abs(tot2[f]))


def test_conservation_sums_one_term_after_another():
    """The running sum is sequential over (element, point): with one field
    holding a huge constant in element 0 and small values elsewhere, the
    small terms are absorbed one by one (their sum, not rounded in order,
    would survive).  Element integrals of a constant c are c*J_e*sum(w)."""
    mesh = _mesh(1, 1, nx=4)
    prob = oracle.Problem(mesh.tables(), MULTIFLUID, ROE)
    NE, Ns = prob.cfg.N_E, prob.cfg.N_s
    V = np.zeros((NE, prob.nf, Ns))
    V[0, 0] = 2.0 ** 53  # element 0 (J = 3/8, weights 1, 1): terms 3/8 * 2^53 each
    V[1:, 0] = 1.0 / 3.0  # later terms of 1/8 each: below half an ulp of the running sum
    tot = prob.conservation(V.ravel())
    J = mesh.tables()["J"]
    w = mesh.tables()["weight"]
    run = 0.0
    for e in range(NE):
        for g in range(w.size):
            run += V[e, 0, 0] * J[e] * w[g]  # every point of a constant element has the node value
    assert tot[0] == run
    assert tot[0] != math.fsum(V[e, 0, 0] * J[e] * w[g] for e in range(NE) for g in range(w.size))
