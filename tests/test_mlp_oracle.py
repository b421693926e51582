"""CPU checks of the indicator's restatement (oracle/mlp.py).

The gradient is pinned by central finite differences of the same loss
(float64, no BF16 rounding); the features are pinned by closed forms: the
tensor-Legendre transform reproduces polynomial data exactly, its mode 0
is the cell mean, and a uniform state has no neighbour jumps.
"""
import numpy as np

import oracle.mlp as om
from dg_amd import MULTIFLUID, PERIODIC, Mesh


def _mesh():
    return Mesh(2, 2, 6, 4, 0.0, 1.5, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = rng.standard_normal(om.N_PARAM) * 0.3
    P[om.W1:om.B1].reshape(om.N_H, om.N_PAD)[:, om.N_IN:] = 0.0
    X = np.zeros((40, om.N_PAD))
    X[:, :om.N_IN] = rng.standard_normal((40, om.N_IN))
    y = (rng.random(40) < 0.3).astype(float)
    g = om.gradient(P, X, y, 3.0, round_bf16=False)
    idx = list(rng.choice(om.N_PARAM, 60, replace=False)) + [om.B1, om.W2 + 5, om.B2 + 7, om.W3 + 3, om.B3]
    for k in idx:
        if om.W1 <= k < om.B1 and (k - om.W1) % om.N_PAD >= om.N_IN:
            continue
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (om.loss(Pp, X, y, 3.0, False) - om.loss(Pm, X, y, 3.0, False)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 0.0], dtype=np.float32)
    r = om.bf16(v)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0 and r[4] == -2.5 and r[5] == 0.0


def test_features_of_polynomial_and_uniform_states():
    mesh = _mesh()
    t = mesh.tables()
    vinv = om.legendre_vinv(t["ref_nodes"])
    xy = t["ref_nodes"].reshape(-1, 2)
    NE = t["N_E"]
    # per element: fields = polynomial in reference coordinates with known modes
    U = np.zeros((NE, 5, 9))
    modes = np.zeros((NE, 5, 9))
    rng = np.random.default_rng(1)
    P = [lambda x: np.ones_like(x), lambda x: x, lambda x: 0.5 * (3 * x * x - 1)]
    for e in range(NE):
        for f in range(5):
            c = rng.standard_normal(9) * 0.1
            c[0] = 2.0 + e * 0.01  # positive means
            modes[e, f] = c
            U[e, f] = sum(c[q] * P[q // 3](xy[:, 0]) * P[q % 3](xy[:, 1]) for q in range(9))
    X = om.features(U.ravel(), t["neighbors"], vinv)
    mean = modes[:, :, 0]
    sc = np.stack([mean[:, 0], np.sqrt(mean[:, 0] * mean[:, 3]), np.sqrt(mean[:, 0] * mean[:, 3]), mean[:, 3],
                   mean[:, 4]], axis=1)
    expect = np.clip(modes / sc[:, :, None], -8, 8).reshape(NE, 45)
    assert np.allclose(X[:, :45], expect, rtol=1e-6, atol=1e-6)
    assert np.all(X[:, 49:] == 0)
    # a uniform state: only mode 0, no jumps
    Uc = np.tile(np.array([1.0, 0.2, 0.1, 3.0, 2.5])[:, None], (NE, 1, 9))
    Xc = om.features(Uc.ravel(), t["neighbors"], vinv)
    assert np.allclose(Xc[:, 45:49], 0.0)
    assert np.allclose(Xc[:, [0, 27, 36]], 1.0)


def test_clamp_after_conversion_equals_clamp_before():
    """The indicator clamps a scaled mode to [-8, 8] after its conversion to
    float (single-precision min/max) where the features of the reference
    form clamp the double first: the conversion is monotonic and +-8 are
    floats, so both give the same float for every double, infinities to +-8,
    NaN to +0 and -0 kept.  Checked over the edges (values that round to +-8,
    the largest doubles, subnormals, zeros, NaN) and random values."""
    def before(v):
        c = 8.0 if v > 8.0 else (-8.0 if v < -8.0 else (v if v == v else 0.0))
        return np.float32(c)

    def after(v):
        x = np.float32(v)
        return np.float32(min(max(x, np.float32(-8.0)), np.float32(8.0))) if x == x else np.float32(0.0)

    eps = np.finfo(np.float64).eps
    edges = [0.0, -0.0, 8.0, -8.0, 8.0 * (1 + eps), -8.0 * (1 + eps), 8.0 * (1 - eps / 2), np.nextafter(8.0, 9.0),
             np.nextafter(8.0, 0.0), np.nextafter(-8.0, 0.0), 8.0000001, -8.0000001, 7.9999999, np.inf, -np.inf,
             np.nan, 5e-324, -5e-324, 1e-310, np.finfo(np.float64).max, -np.finfo(np.float64).max, 3.4e38, 1e39]
    rng = np.random.default_rng(7)
    vals = edges + list(rng.standard_normal(20000) * 6.0) + list(rng.standard_normal(2000) * 1e-30)
    with np.errstate(over="ignore", invalid="ignore"):
        for v in vals:
            b, a = before(v), after(v)
            assert b.tobytes() == a.tobytes(), (v, b, a)
