"""Shared helpers for the tests: problem construction and comparisons."""
from __future__ import annotations

import numpy as np

import dg_amd
from dg_amd import (FARFIELD, MULTIFLUID, PASSIVE, PERIODIC, REFLECTIVE, ROE, RUSANOV, SINGLEFLUID, STIFFENED, Config,
                    Mesh, nfields)
from dg_amd import problems
import oracle

CONFIGS = [(d, ph, fl) for d in (1, 2) for ph in (SINGLEFLUID, MULTIFLUID) for fl in (RUSANOV, ROE)]
# passive-scalar and stiffened-gas configurations (face + element kernels)
CONFIGS_PS = [(d, ph, fl) for d in (1, 2) for ph in (PASSIVE, STIFFENED) for fl in (RUSANOV, ROE)]


def cfg_id(c):
    d, ph, fl = c
    return oracle.ref_name(d, ph, fl)


def small_mesh(dim, order, physics, bc=None, n=(12, 7)):
    nf = nfields(dim, physics)
    if dim == 1:
        bc = bc or (FARFIELD, REFLECTIVE)
        return Mesh(1, order, n[0], 1, -1.0, 1.0, bc=bc, nfields=nf)
    bc = bc or (PERIODIC, PERIODIC, REFLECTIVE, FARFIELD)
    return Mesh(2, order, n[0], n[1], 0.0, 1.2, 0.0, 0.7, bc=bc, nfields=nf)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def per_field_rel(U, Uref, nf, ns):
    """Relative L2 and Linf error of each conserved field."""
    A = np.asarray(U).reshape(-1, nf, ns)
    B = np.asarray(Uref).reshape(-1, nf, ns)
    out = []
    for f in range(nf):
        d = (A[:, f] - B[:, f]).ravel()
        r = B[:, f].ravel()
        l2 = np.linalg.norm(d) / max(np.linalg.norm(r), 1e-300)
        li = np.abs(d).max() / max(np.abs(r).max(), 1e-300)
        out.append((l2, li))
    return out
