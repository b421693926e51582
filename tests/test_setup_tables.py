"""Setup tables (reference main.cc:312-740, dg_functions.cc, simpleMesh.cc).

The reference's setup code needs LAPACK (fullMatrix::invert), which the image
lacks, so these tables are pinned by what the mathematics fixes: Gmsh node and
closure conventions (polynomialBasis.cc:20-31, 713-751, 911-934), Gauss rules
by exactness, the Lagrange property, closed-form Jacobians and mass matrices
of rectangles, geometric consistency of faces/normals/maps/neighbours, and an
independent numpy restatement of the reference's Lagrange->Taylor
construction (Vandermonde at the Gauss points, inverted).
"""
import math

import numpy as np
import pytest

from dg_amd import FARFIELD, PERIODIC, REFLECTIVE, Mesh

GMSH_NODES = {
    (1, 1): [(-1,), (1,)],
    (1, 2): [(-1,), (1,), (0,)],
    (1, 3): [(-1,), (1,), (-1 / 3,), (1 / 3,)],
    (2, 1): [(-1, -1), (1, -1), (1, 1), (-1, 1)],
    (2, 2): [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)],
    (2, 3): [(-1, -1), (1, -1), (1, 1), (-1, 1), (-1 / 3, -1), (1 / 3, -1), (1, -1 / 3), (1, 1 / 3), (1 / 3, 1),
             (-1 / 3, 1), (-1, 1 / 3), (-1, -1 / 3), (-1 / 3, -1 / 3), (1 / 3, -1 / 3), (1 / 3, 1 / 3),
             (-1 / 3, 1 / 3)],
}


def mesh_for(dim, order, bc=None, n=(5, 4), ext=(0.0, 1.5, -0.5, 0.3)):
    if dim == 1:
        return Mesh(1, order, n[0], 1, ext[0], ext[1], bc=bc or (FARFIELD, REFLECTIVE), nfields=4)
    return Mesh(2, order, n[0], n[1], *ext, bc=bc or (PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), nfields=5)


def basis_eval(dim, order, nodes, pts):
    """Lagrange basis on the tensor of 1D Gmsh line nodes, evaluated at pts."""
    ln = np.array([x[0] for x in GMSH_NODES[(1, order)]])

    def L(k, x):
        v = np.ones_like(x)
        for m in range(len(ln)):
            if m != k:
                v = v * (x - ln[m]) / (ln[k] - ln[m])
        return v

    out = np.zeros((len(pts), len(nodes)))
    for i, nd in enumerate(nodes):
        a = int(np.argmin(np.abs(ln - nd[0])))
        v = L(a, pts[:, 0])
        if dim == 2:
            b = int(np.argmin(np.abs(ln - nd[1])))
            v = v * L(b, pts[:, 1])
        out[:, i] = v
    return out


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3)])
def test_reference_element(dim, order):
    m = mesh_for(dim, order)
    s = m.sizes
    nodes = m.ref_nodes.reshape(-1, dim)
    assert np.allclose(nodes, np.array(GMSH_NODES[(dim, order)]), atol=1e-15)
    # quadrature: n points exact to degree 2n-1 per direction (GaussQuadratureQuad.cc:113-149)
    pts, w = m.points.reshape(-1, dim), m.weight
    n1 = order + 1 if dim == 1 else (2 if order == 1 else order + 2)
    for a in range(2 * n1):
        exact = 0.0 if a % 2 else 2.0 / (a + 1)
        if dim == 1:
            assert abs(np.sum(w * pts[:, 0] ** a) - exact) < 1e-14
        else:
            for b in range(2 * n1):
                eb = 0.0 if b % 2 else 2.0 / (b + 1)
                assert abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b) - exact * eb) < 1e-14
    assert len(w) == {1: order + 1, 2: (4 if order == 1 else (order + 2) ** 2)}[dim]
    if dim == 2 and order >= 2:  # x is the slow index (GaussQuadratureQuad.cc:128-133)
        n = order + 2
        assert np.all(np.diff(pts[:, 0].reshape(n, n)[:, 0]) > 0) and np.all(np.ptp(pts[:, 0].reshape(n, n), axis=1) == 0)
    # Lagrange property and phi/dphi tables (main.cc:372-413)
    NG, Ns = s["N_G"], s["N_s"]
    phi = m.phi.reshape(Ns, NG).T
    assert np.allclose(basis_eval(dim, order, nodes, nodes), np.eye(Ns), atol=1e-14)
    assert np.allclose(phi, basis_eval(dim, order, nodes, pts), atol=1e-14)
    assert np.allclose(phi.sum(1), 1.0, atol=1e-14)
    dphi = m.dphi.reshape(Ns, NG * dim).T
    h = 1e-6
    for al in range(dim):
        dp = pts.copy()
        dm = pts.copy()
        dp[:, al] += h
        dm[:, al] -= h
        fd = (basis_eval(dim, order, nodes, dp) - basis_eval(dim, order, nodes, dm)) / (2 * h)
        assert np.allclose(dphi[al::dim], fd, atol=1e-8)
        assert np.allclose(dphi[al::dim].sum(1), 0.0, atol=1e-13)
    assert np.allclose(m.phi_w.reshape(Ns, NG).T, phi * w[:, None], atol=0)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_closures(order):
    m = mesh_for(2, order)
    Ms = order + 1
    cl = m.closures.reshape(8, Ms)
    nodes = m.ref_nodes.reshape(-1, 2)
    corners = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    for j in range(4):
        a, b = np.array(corners[j]), np.array(corners[(j + 1) % 4])
        pts = nodes[cl[j]]
        # endpoints then interior nodes walking from a to b (polynomialBasis.cc:911-926)
        assert np.allclose(pts[0], a) and np.allclose(pts[1], b)
        for k in range(2, Ms):
            assert np.allclose(pts[k], a + (k - 1) * (b - a) / order)
        rev = nodes[cl[4 + j]]
        assert np.allclose(rev[0], b) and np.allclose(rev[1], a)
        assert np.allclose(rev[2:], pts[2:][::-1])


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3)])
@pytest.mark.parametrize("bcset", [0, 1])
def test_faces_normals_maps(dim, order, bcset):
    if dim == 1:
        bc = [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC)][bcset]
    else:
        bc = [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (REFLECTIVE, FARFIELD, PERIODIC, PERIODIC)][bcset]
    ext = (0.0, 1.5, -0.5, 0.3)
    m = mesh_for(dim, order, bc=bc, ext=ext)
    s = m.sizes
    Ns, Ms, MT, NE, NF, NN = s["N_s"], s["M_s"], s["M_T"], s["N_E"], s["N_F"], s["N_N"]
    xyz = m.xyz.reshape(NE, Ns, dim)
    fe, fc, tag = m.face_elem.reshape(MT, 2), m.face_closure.reshape(MT, 2), m.face_tag
    cl = m.closures.reshape(-1, Ms)
    nrm = m.normals.reshape(MT, dim)
    L = np.array([ext[1] - ext[0], ext[3] - ext[2]])[:dim]
    cen = xyz.mean(1)
    for t in range(MT):
        p0 = xyz[fe[t, 0], cl[fc[t, 0]]]
        p1 = xyz[fe[t, 1], cl[fc[t, 1]]]
        d = p1 - p0
        if tag[t] == PERIODIC:
            d = d - np.round(d / L) * L  # periodic images coincide
        assert np.allclose(d, 0, atol=1e-12), t
        # unit normal, outward from side 0
        assert abs(np.linalg.norm(nrm[t]) - 1) < 1e-15
        assert np.dot(p0.mean(0) - cen[fe[t, 0]], nrm[t]) > 0
        if tag[t] in (FARFIELD, REFLECTIVE):
            assert fe[t, 0] == fe[t, 1] and fc[t, 0] == fc[t, 1]
    # map: ((t*NF+f)*2+d)*Ms+j -> (e*NF+f)*Ns + cl[j]
    mp = m.map.reshape(MT, NF, 2, Ms)
    for f in range(NF):
        for dd in range(2):
            exp = (fe[:, dd][:, None] * NF + f) * Ns + cl[fc[:, dd]]
            assert np.array_equal(mp[:, f, dd], exp)
    # invmap: each (e,f) lists M_s*N_N (node, Qtcj) pairs covering its face nodes
    inv = m.invmap.reshape(NE, NF, Ms * NN, 2)
    for e in range(NE):
        nodes_hit = np.bincount(inv[e, 0, :, 0], minlength=Ns)
        corners = 2 if dim == 2 else 1
        expect = np.zeros(Ns, int)
        for k in range(NN):
            for j in range(Ms):
                expect[cl[k, j]] += 1
        assert np.array_equal(nodes_hit, expect)
        q = inv[e, 0, :, 1]
        t_of = q // (NF * 2 * Ms)
        side = (q // Ms) % 2
        assert np.all(fe[t_of, side] == e)
    # neighbours LRDU (simpleMesh.cc:664-754)
    nb = m.neighbors.reshape(NE, NN)
    h = (np.array([ext[1] - ext[0], ext[3] - ext[2]])[:dim]) / np.array([s["nx"], s["ny"]])[:dim]
    for e in range(NE):
        for k in range(NN):
            o = nb[e, k]
            axis, sgn = k // 2, (-1 if k % 2 == 0 else 1)
            if o < 0:
                side = 2 * axis + (0 if sgn < 0 else 1)
                assert -o == bc[side] and bc[side] != PERIODIC
            else:
                dc = cen[o] - cen[e]
                dc = dc - np.round(dc / L) * L
                exp = np.zeros(dim)
                exp[axis] = sgn * h[axis]
                assert np.allclose(dc, exp, atol=1e-12)
    # boundary map lists the wall faces
    assert np.array_equal(m.boundaryMap, np.nonzero(tag == REFLECTIVE)[0])


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3)])
def test_geometry_and_mass(dim, order):
    ext = (0.0, 1.5, -0.5, 0.3)
    m = mesh_for(dim, order, ext=ext)
    s = m.sizes
    NE, NG, Ns = s["N_E"], s["N_G"], s["N_s"]
    hx = (ext[1] - ext[0]) / s["nx"]
    hy = (ext[3] - ext[2]) / s["ny"] if dim == 2 else None
    J = m.J
    if dim == 1:
        assert np.allclose(J, hx / 2, rtol=1e-14)
        assert np.allclose(m.invJac, 2 / hx, rtol=1e-14)
        assert np.allclose(m.JF, 1.0)
        assert abs(m.dx - hx) < 1e-14
    else:
        assert np.allclose(J, hx * hy / 4, rtol=1e-14)
        ij = m.invJac.reshape(NE, NG, 2, 2)
        assert np.allclose(ij[..., 0, 0], 2 / hx) and np.allclose(ij[..., 1, 1], 2 / hy)
        assert np.allclose(ij[..., 0, 1], 0) and np.allclose(ij[..., 1, 0], 0)
        nrm = m.normals.reshape(-1, 2)
        jf = m.JF.reshape(-1, 2)
        horiz = np.abs(nrm[:, 1]) > 0.5
        assert np.allclose(jf[horiz], hx / 2) and np.allclose(jf[~horiz], hy / 2)
        assert abs(m.dx - min(hx, hy)) < 1e-14
    # Minv_e * M_e = I with M_e = J_e * int phi_i phi_k (exact high-order rule)
    nodes = m.ref_nodes.reshape(-1, dim)
    x, w = np.polynomial.legendre.leggauss(8)
    if dim == 1:
        pts, ww = x[:, None], w
    else:
        X, Y = np.meshgrid(x, x, indexing="ij")
        pts, ww = np.stack([X.ravel(), Y.ravel()], 1), np.outer(w, w).ravel()
    B = basis_eval(dim, order, nodes, pts)
    Mref = B.T @ (B * ww[:, None])
    Minv = m.Minv.reshape(NE, Ns, Ns)  # [e, k, i] = Minv_e(i, k)
    for e in (0, NE - 1):
        assert np.allclose(Minv[e].T @ (J[e] * Mref), np.eye(Ns), atol=1e-10)


def _ref_lagrange_to_taylor(order):
    """Restates main.cc:557-567 / 1471-1506: Transform1D = inv(monoV) * phi at
    the p+1 Gauss points (monoV(g,j) = x_g^j / j!)."""
    x, _ = np.polynomial.legendre.leggauss(order + 1)
    ln = np.array([v[0] for v in GMSH_NODES[(1, order)]])
    V = np.array([[xg ** j / math.factorial(j) for j in range(order + 1)] for xg in x])
    phi = basis_eval(1, order, [(v,) for v in ln], x[:, None])
    return np.linalg.inv(V) @ phi


@pytest.mark.parametrize("order", [1, 2, 3])
def test_limiter_transforms(order):
    T = _ref_lagrange_to_taylor(order)
    m1 = mesh_for(1, order)
    n = order + 1
    assert np.allclose(m1.Lag2Mono.reshape(n, n).T, T, atol=1e-13)
    assert np.allclose(m1.Mono2Lag.reshape(n, n).T, np.linalg.inv(T), atol=1e-13)
    # 2D (main.cc:1348-1524): x-slices at y-levels (min, max, interior), y-slices at x-levels
    m2 = mesh_for(2, order)
    Ns = n * n
    L2MX = m2.Lag2MonoX.reshape(Ns, Ns).T
    MX2MY = m2.MonoX2MonoY.reshape(Ns, Ns).T
    MY2L = m2.MonoY2Lag.reshape(Ns, Ns).T
    L2MY = m2.Lag2MonoY.reshape(Ns, Ns).T
    MX2L = m2.MonoX2Lag.reshape(Ns, Ns).T
    nodes = m2.ref_nodes.reshape(-1, 2)
    lv = np.array([v[0] for v in GMSH_NODES[(1, order)]])
    for a in range(n):
        for b in range(n):
            u = nodes[:, 0] ** a / math.factorial(a) * nodes[:, 1] ** b / math.factorial(b)
            ax = (L2MX @ u).reshape(n, n)  # [y-level slice, x-power]
            ay = (MX2MY @ (L2MX @ u)).reshape(n, n)  # [x-level slice, y-power]
            for sl in range(n):
                ex = np.zeros(n)
                ex[a] = lv[sl] ** b / math.factorial(b)
                assert np.allclose(ax[sl], ex, atol=1e-13)
                ey = np.zeros(n)
                ey[b] = lv[sl] ** a / math.factorial(a)
                assert np.allclose(ay[sl], ey, atol=1e-13)
            assert np.allclose(MY2L @ (L2MY @ u), u, atol=1e-13)
            assert np.allclose(MX2L @ (L2MX @ u), u, atol=1e-13)


def _header_tables():
    """The constant transforms compiled into the HR kernels (hr_tables.h),
    read from the header's initialisers."""
    import os
    import re

    path = os.path.join(os.path.dirname(__file__), "..", "dg_amd", "csrc", "hr_tables.h")
    text = open(path).read()
    out = {}
    for p, body in re.findall(r"struct HrTables2D<(\d)> \{(.*?)\n\};", text, re.S):
        for name, vals in re.findall(r"static constexpr double (\w+)\[\d+\] = \{(.*?)\};", body, re.S):
            out[(int(p), name)] = np.array([float(v) for v in vals.replace("\n", " ").split(",")])
    return out


@pytest.mark.parametrize("order", [1, 2])
def test_constant_hr_transforms_equal_the_built_tables(order):
    """The HR kernels carry Lag2MonoX, MonoX2MonoY and MonoY2Lag as constants;
    the solver compares the built tables with them before it uses those
    kernels.  They must be the built tables exactly (the mesh-independent
    reference-element transforms), or the constant-table kernels would go
    unused."""
    tabs = _header_tables()
    for n, ext in (((5, 4), (0.0, 1.5, -0.5, 0.3)), ((3, 7), (-2.0, 1.0, 0.0, 4.0))):
        t = mesh_for(2, order, n=n, ext=ext).tables()
        for key, name in (("Lag2MonoX", "L2MX"), ("MonoX2MonoY", "MX2MY"), ("MonoY2Lag", "MY2L")):
            built = np.asarray(t[key], dtype=np.float64).ravel()
            const = tabs[(order, name)]
            assert built.shape == const.shape
            assert np.array_equal(built, const), (key, order)
            # zero entries are exact zeros (the kernels leave those terms out)
            assert np.array_equal(built == 0.0, const == 0.0)
