"""Initial states for the benchmark configurations and the parity tests.

Like the reference's initial conditions (init_cond.cc), states are set by
nodal interpolation at the mesh nodes and returned in the reference layout
U[(e*N_F + f)*N_s + i].  Multifluid states carry alpha = 1/(gamma-1)
(GAMNCON).
"""
from __future__ import annotations

import math

import numpy as np

from . import FARFIELD, MULTIFLUID, PASSIVE, PERIODIC, REFLECTIVE, ROE, RUSANOV, SINGLEFLUID, STIFFENED, Config, Mesh, nfields

GAMMA = 1.4


def _pack(mesh: Mesh, fields: list[np.ndarray]) -> np.ndarray:
    """fields[f] has shape (N_E, N_s) -> U[(e*NF+f)*N_s+i]; one field per
    field of the mesh."""
    if len(fields) != mesh.sizes["N_F"]:
        raise ValueError(f"{len(fields)} fields for a mesh of {mesh.sizes['N_F']}")
    return np.ascontiguousarray(np.stack(fields, axis=1).reshape(-1))


def node_xy(mesh: Mesh):
    s = mesh.sizes
    xyz = mesh.array("xyz").reshape(s["N_E"], s["N_s"], s["D"])
    x = xyz[..., 0]
    y = xyz[..., 1] if s["D"] == 2 else None
    return x, y


def _closure(physics, rho, p, gam, ke, pinf=None, phic=None, phinc=None):
    """[E, extra fields...] for the physics: E = p/(gamma-1) + ke (+ gamma
    pinf/(gamma-1), stiffened); two-fluid adds 1/(gamma-1); passive adds
    rho*phic and phinc; stiffened adds 1/(gamma-1) and gamma pinf/(gamma-1)."""
    if physics == STIFFENED:
        beta = gam * pinf / (gam - 1.0)
        return [p / (gam - 1.0) + beta + ke, 1.0 / (gam - 1.0), beta]
    f = [p / (gam - 1.0) + ke]
    if physics == MULTIFLUID:
        f.append(1.0 / (gam - 1.0))
    elif physics == PASSIVE:
        f += [rho * phic, phinc]
    return f


def conserved_1d(rho, u, p, gamma, physics):
    E = p / (gamma - 1.0) + 0.5 * rho * u * u
    f = [rho, rho * u, E]
    if physics == MULTIFLUID:
        f.append(np.broadcast_to(1.0 / (gamma - 1.0), np.shape(rho)).astype(float))
    return f


def conserved_2d(rho, u, v, p, gamma, physics):
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    f = [rho, rho * u, rho * v, E]
    if physics == MULTIFLUID:
        f.append(np.broadcast_to(1.0 / (gamma - 1.0), np.shape(rho)).astype(float))
    return f


# ----------------------------------------------------------------- 1D Sod
def sod_mesh(nx: int = 1000, order: int = 2, physics: int = MULTIFLUID) -> Mesh:
    return Mesh(1, order, nx, 1, -1.0, 1.0, bc=(FARFIELD, FARFIELD), nfields=nfields(1, physics))


def sod_state(mesh: Mesh, physics: int = MULTIFLUID) -> np.ndarray:
    """Sod tube, gamma=1.4 both sides, split on each element's first node
    (buildLRstates_multifluid, init_cond.cc:351-397).  Passive scalars: 1 on
    the left, 0 on the right; stiffened gas: pinf = 0."""
    x, _ = node_xy(mesh)
    left = (x[:, :1] < 1e-8) & np.ones_like(x, dtype=bool)
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    u = np.zeros_like(x)
    if physics in (SINGLEFLUID, MULTIFLUID):
        return _pack(mesh, conserved_1d(rho, u, p, GAMMA, physics))
    phi = np.where(left, 1.0, 0.0)
    extra = dict(phic=phi, phinc=phi) if physics == PASSIVE else dict(pinf=np.zeros_like(x))
    return _pack(mesh, [rho, rho * u] + _closure(physics, rho, p, np.full_like(x, GAMMA), 0.5 * rho * u * u, **extra))


# -------------------------------------------------------- 2D Shu vortex
def vortex_mesh(n: int = 512, order: int = 2) -> Mesh:
    return Mesh(2, order, n, n, 0.0, 10.0, 0.0, 10.0, bc=(PERIODIC,) * 4, nfields=4)


def vortex_fields(x, y, t: float = 0.0, beta: float = 5.0, xc: float = 5.0, yc: float = 5.0, L: float = 10.0):
    """Isentropic vortex in a (1,1,1,1) free stream, advected to time t on a
    periodic [0,L]^2 (exact solution)."""
    g = GAMMA
    dx = np.mod(x - xc - t + L / 2, L) - L / 2
    dy = np.mod(y - yc - t + L / 2, L) - L / 2
    r2 = dx * dx + dy * dy
    ex = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - beta / (2 * math.pi) * ex * dy
    v = 1.0 + beta / (2 * math.pi) * ex * dx
    T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * math.pi * math.pi) * ex * ex
    rho = T ** (1.0 / (g - 1.0))
    p = rho ** g
    return rho, u, v, p


def vortex_state(mesh: Mesh, t: float = 0.0) -> np.ndarray:
    x, y = node_xy(mesh)
    rho, u, v, p = vortex_fields(x, y, t)
    return _pack(mesh, conserved_2d(rho, u, v, p, GAMMA, SINGLEFLUID))


# ------------------------------------------------- 2D shock-bubble (5 fields)
def shock_bubble_mesh(nx: int = 1024, ny: int = 512, order: int = 2) -> Mesh:
    return Mesh(2, order, nx, ny, 0.0, 2.0, 0.0, 1.0, bc=(FARFIELD, FARFIELD, REFLECTIVE, REFLECTIVE), nfields=5)


def shock_bubble_state(mesh: Mesh, mach: float = 1.22, xs: float = 0.1, xb: float = 0.5, yb: float = 0.5,
                       rb: float = 0.25, rho_b: float = 0.138, gamma_b: float = 1.667) -> np.ndarray:
    """Planar shock at x=xs moving right into quiescent air (rho=1, p=1,
    gamma=1.4) with a light-gas bubble in pressure equilibrium; post-shock
    state from the Rankine-Hugoniot relations."""
    g = GAMMA
    x, y = node_xy(mesh)
    M2 = mach * mach
    p2 = 1.0 + 2 * g / (g + 1) * (M2 - 1.0)
    rho2 = (g + 1) * M2 / ((g - 1) * M2 + 2.0)
    c1 = math.sqrt(g)
    s = mach * c1
    u2 = s * (1.0 - 1.0 / rho2)
    # cell-based split on each element's centre, like the reference's per-element ICs
    xc = x.mean(axis=1, keepdims=True) * np.ones_like(x)
    yc = y.mean(axis=1, keepdims=True) * np.ones_like(y)
    shocked = xc < xs
    bubble = (xc - xb) ** 2 + (yc - yb) ** 2 < rb * rb
    rho = np.where(shocked, rho2, np.where(bubble, rho_b, 1.0))
    u = np.where(shocked, u2, 0.0)
    v = np.zeros_like(x)
    p = np.where(shocked, p2, 1.0)
    gam = np.where(bubble & ~shocked, gamma_b, g)
    E = p / (gam - 1.0) + 0.5 * rho * (u * u + v * v)
    return _pack(mesh, [rho, rho * u, rho * v, E, 1.0 / (gam - 1.0)])


# ------------------------------------------------------- parity test states
def smooth_random_state(mesh: Mesh, physics: int, seed: int = 0, amp: float = 0.2) -> np.ndarray:
    """Seeded smooth positive states with O(amp) variations (all fields vary)."""
    rng = np.random.default_rng(seed)
    x, y = node_xy(mesh)
    D = mesh.sizes["D"]
    x0, x1, y0, y1 = mesh.extent
    X = (x - x0) / (x1 - x0) * 2 * math.pi
    Y = ((y - y0) / (y1 - y0) * 2 * math.pi) if D == 2 else np.zeros_like(x)

    def wave():
        a, b, c, d = rng.uniform(-1, 1, 4)
        return a * np.sin(X + 2 * math.pi * c) + b * np.cos(Y + 2 * math.pi * d) + 0.3 * np.sin(2 * X + Y)

    rho = 1.0 + amp * wave()
    p = 1.0 + amp * wave()
    u = 0.5 * amp * wave() + rng.uniform(-0.3, 0.3)
    gam = (1.4 + 0.2 * (0.5 + 0.5 * np.tanh(wave())) if physics in (MULTIFLUID, STIFFENED)
           else np.full_like(x, GAMMA))
    extra = {}
    if physics == STIFFENED:
        extra["pinf"] = 0.5 + 0.5 * amp * wave()
    elif physics == PASSIVE:
        extra = {"phic": 0.5 + amp * wave(), "phinc": 0.5 + amp * wave()}
    if D == 1:
        f = [rho, rho * u] + _closure(physics, rho, p, gam, 0.5 * rho * u * u, **extra)
    else:
        v = 0.5 * amp * wave() + rng.uniform(-0.3, 0.3)
        f = [rho, rho * u, rho * v] + _closure(physics, rho, p, gam, 0.5 * rho * (u * u + v * v), **extra)
    return _pack(mesh, f)


def _side_extras(physics, rng, left):
    """Per-side pinf (stiffened) or phic/phinc (passive) for the Riemann states."""
    if physics == STIFFENED:
        a, b = rng.uniform(0.0, 2.0, 2)
        return {"pinf": np.where(left, a, b)}
    if physics == PASSIVE:
        c = rng.uniform(0.0, 1.0, 4)
        return {"phic": np.where(left, c[0], c[1]), "phinc": np.where(left, c[2], c[3])}
    return {}


def riemann_state(mesh: Mesh, physics: int, seed: int = 0, x_split: float | None = None) -> np.ndarray:
    """Seeded Riemann problem: rho, p log-uniform in [0.1, 10], u in [-2, 2],
    gamma in [1.1, 1.7] per side (multifluid), split on element first node."""
    rng = np.random.default_rng(seed)
    x, y = node_xy(mesh)
    D = mesh.sizes["D"]
    x0, x1, _, _ = mesh.extent
    xs = 0.5 * (x0 + x1) if x_split is None else x_split
    left = (x[:, :1] < xs) & np.ones_like(x, dtype=bool)
    rl, rr = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    pl, pr = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    ul, ur = rng.uniform(-2, 2, 2)
    gl, gr = (rng.uniform(1.1, 1.7, 2) if physics in (MULTIFLUID, STIFFENED) else (GAMMA, GAMMA))
    rho = np.where(left, rl, rr)
    p = np.where(left, pl, pr)
    u = np.where(left, ul, ur)
    gam = np.where(left, gl, gr)
    extra = _side_extras(physics, rng, left)
    tail = _closure(physics, rho, p, gam, 0.5 * rho * u * u, **extra)
    f = [rho, rho * u] + ([] if D == 1 else [np.zeros_like(x)]) + tail
    return _pack(mesh, f)


def oblique_riemann_state(mesh: Mesh, physics: int, seed: int = 0) -> np.ndarray:
    """Seeded 2D Riemann problem across an oblique line through the domain
    centre (random angle): rho, p log-uniform in [0.5, 2], u and v in
    [-0.5, 0.5], gamma in [1.1, 1.7] per side (multifluid).  Every field,
    both momenta included, varies, and elements are constant (the side is
    chosen by the element's first node)."""
    rng = np.random.default_rng(seed)
    x, y = node_xy(mesh)
    x0, x1, y0, y1 = mesh.extent
    th = rng.uniform(0.2, 1.3)
    side = (x[:, :1] - 0.5 * (x0 + x1)) * math.cos(th) + (y[:, :1] - 0.5 * (y0 + y1)) * math.sin(th)
    left = (side < 0) & np.ones_like(x, dtype=bool)
    rl, rr = np.exp(rng.uniform(np.log(0.5), np.log(2.0), 2))
    pl, pr = np.exp(rng.uniform(np.log(0.5), np.log(2.0), 2))
    ul, ur = rng.uniform(-0.5, 0.5, 2)
    vl, vr = rng.uniform(-0.5, 0.5, 2)
    gl, gr = (rng.uniform(1.1, 1.7, 2) if physics in (MULTIFLUID, STIFFENED) else (GAMMA, GAMMA))
    rho = np.where(left, rl, rr)
    p = np.where(left, pl, pr)
    u = np.where(left, ul, ur)
    v = np.where(left, vl, vr)
    gam = np.where(left, gl, gr)
    extra = _side_extras(physics, rng, left)
    f = [rho, rho * u, rho * v] + _closure(physics, rho, p, gam, 0.5 * rho * (u * u + v * v), **extra)
    return _pack(mesh, f)


# ------------------------------------------------ batched 1D ensemble
def ensemble_spec(n: int = 256, seed: int = 20261202, ne_range=(200, 4000), orders=(1, 2, 3)):
    """Seeded sizes of an ensemble of 1D problems: element counts uniform in
    ne_range, orders uniform in `orders`."""
    rng = np.random.default_rng(seed)
    ne = rng.integers(ne_range[0], ne_range[1] + 1, n)
    p = rng.choice(np.asarray(orders), n)
    return [int(x) for x in ne], [int(x) for x in p]


def two_fluid_riemann_1d(mesh: Mesh, seed: int) -> np.ndarray:
    """Seeded 1D two-fluid Riemann problem: rho, p log-uniform in [0.1, 10],
    u uniform in [-2, 2], gamma uniform in [1.1, 1.7] per side, split at x=0
    on each element's first node (4 fields: rho, rho u, E, 1/(gamma-1))."""
    rng = np.random.default_rng(seed)
    x, _ = node_xy(mesh)
    left = (x[:, :1] < 1e-12) & np.ones_like(x, dtype=bool)
    rl, rr = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    pl, pr = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    ul, ur = rng.uniform(-2.0, 2.0, 2)
    gl, gr = rng.uniform(1.1, 1.7, 2)
    rho = np.where(left, rl, rr)
    p = np.where(left, pl, pr)
    u = np.where(left, ul, ur)
    gam = np.where(left, gl, gr)
    E = p / (gam - 1.0) + 0.5 * rho * u * u
    return _pack(mesh, [rho, rho * u, E, 1.0 / (gam - 1.0)])
