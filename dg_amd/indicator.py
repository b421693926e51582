"""Training data, training and evaluation of the learned troubled-cell indicator.

The data come from the solver itself: a mosaic mesh of 2D p=2 two-fluid
cells holds one seeded random Riemann problem per sub-square (half planar,
jump normal to x or y, half oblique or circular), is advanced a few RK4
steps with the global HR limiter on the GPU, and the reference sensor
(calc_sensors semantics, `Solver.sensing`) labels every cell at a few
snapshots.  Only cells away from the sub-square borders are kept, so the
artificial jumps between neighbouring problems never enter the data.
Random multimode interfaces (some hit by a shock) and smooth random flows,
labelled the same way, complete the data.
Features come from `Solver.features`; training runs on the GPU kernels of
dg_amd/csrc/mlp.hip (`Indicator.train`).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import (FARFIELD, LIMIT_HR, MULTIFLUID, PERIODIC, ROE, Config, DgxError, Indicator, Mesh, Solver, problems)

SENSOR_THRESHOLDS = (0.01, 0.01, 0.01)  # "#sensor thresholds" of the data and of the judge
SEED = 20261202


def _advance(solver: Solver, U, dt_cfl: float, nsteps: int, limiter: int) -> bool:
    """nsteps CFL steps of U; False once the state has lost positivity: a NaN
    time step (the driver's stop) or a non-finite value anywhere (the time
    step passes over NaN speeds after the first node, as the reference's
    idamax does, so a cell that lost positivity need not stop the run)."""
    import torch

    try:
        solver.rk4_integrate(U, 1, dt_cfl, nsteps=nsteps, tf=1e30, limiter=limiter)
    except DgxError:
        return False
    return bool(torch.isfinite(U).all().item())


@dataclass
class Dataset:
    X: np.ndarray       # float32 [n, 64] features
    y: np.ndarray       # float32 [n] 1 = flagged by the sensor
    sensor: np.ndarray  # int32 [n] sensor value (1 contact, 2 shock, 3 gamma contact)


def mosaic_state(mesh: Mesh, n_sub: int, sub: int, seed: int) -> np.ndarray:
    """One Riemann problem per sub x sub block of elements (n_sub^2 problems):
    rho, p log-uniform in [0.1, 10], velocities uniform in [-2, 2], gamma
    uniform in [1.1, 1.7] on each side.  Even problem index: a planar jump
    through the block centre, normal to x or y, with velocity normal to the
    jump only (a 1D problem); odd: an oblique line (2 in 3) or a circle
    through/around the centre, 2D velocities.  The side of each element is
    decided by its first node, so elements are constant."""
    rng = np.random.default_rng(seed)
    K = n_sub * n_sub
    off = rng.uniform(-1.0, 1.0, (K, 2))
    rho = np.exp(rng.uniform(math.log(0.1), math.log(10.0), (K, 2)))
    p = np.exp(rng.uniform(math.log(0.1), math.log(10.0), (K, 2)))
    vel = rng.uniform(-2.0, 2.0, (K, 2, 2))
    gam = rng.uniform(1.1, 1.7, (K, 2))
    kind = rng.integers(0, 6, K)
    th = rng.uniform(0.0, math.pi, K)
    rfrac = rng.uniform(0.2, 0.4, K)

    x, y = problems.node_xy(mesh)
    s = mesh.sizes
    nx = s["nx"]
    e = np.arange(s["N_E"])
    ex, ey = e % nx, e // nx
    h = (mesh.extent[1] - mesh.extent[0]) / nx
    k = (ey // sub) * n_sub + ex // sub
    cx = ((ex // sub) * sub + sub / 2) * h + off[k, 0] * h
    cy = ((ey // sub) * sub + sub / 2) * h + off[k, 1] * h
    dx, dy = x[:, 0] - cx, y[:, 0] - cy
    planar = k % 2 == 0
    normal_x = kind[k] % 2 == 0
    oblique = kind[k] < 4
    side_planar = np.where(normal_x, dx < 0, dy < 0)
    side_oblique = dx * np.cos(th[k]) + dy * np.sin(th[k]) < 0
    side_circle = dx * dx + dy * dy < (rfrac[k] * sub * h) ** 2
    left = np.where(planar, side_planar, np.where(oblique, side_oblique, side_circle))
    j = np.where(left, 0, 1)
    r_, p_, g_ = rho[k, j], p[k, j], gam[k, j]
    u_, v_ = vel[k, j, 0], vel[k, j, 1]
    u_ = np.where(planar & ~normal_x, 0.0, u_)
    v_ = np.where(planar & normal_x, 0.0, v_)
    E = p_ / (g_ - 1.0) + 0.5 * r_ * (u_ * u_ + v_ * v_)
    f = np.stack([r_, r_ * u_, r_ * v_, E, 1.0 / (g_ - 1.0)], axis=1)
    return np.ascontiguousarray(np.repeat(f[:, :, None], s["N_s"], axis=2).reshape(-1))


def generate(n_sub: int = 64, sub: int = 8, snapshots=(0, 2, 4, 6), margin: int = 2, seed: int = SEED,
             cfl: float = 0.3, device="cuda") -> Dataset:
    """Features and sensor labels of the mosaic (n_sub^2 problems) at the
    given step counts, interior cells only."""
    import torch

    n = n_sub * sub
    mesh = Mesh(2, 2, n, n, 0.0, float(n_sub), 0.0, float(n_sub), bc=(FARFIELD,) * 4, nfields=5)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(SENSOR_THRESHOLDS)
    solver.set_indicator(Indicator(seed=0), 0.0)  # allocates the feature scratch
    U = torch.from_numpy(mosaic_state(mesh, n_sub, sub, seed)).to(device)
    NE = mesh.sizes["N_E"]
    e = np.arange(NE)
    ex, ey = e % n, e // n
    keep = ((ex % sub) >= margin) & ((ex % sub) < sub - margin) & ((ey % sub) >= margin) & ((ey % sub) < sub - margin)
    X = torch.zeros((NE, 64), dtype=torch.float32, device=device)
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    cfl_dt = cfl * mesh.dx / 5
    out_X, out_s = [], []
    done = 0
    for snap in snapshots:
        if snap > done:
            if not _advance(solver, U, cfl_dt, snap - done, LIMIT_HR):
                break  # a problem lost positivity: keep the snapshots taken so far
            done = snap
        solver.features(U, X)
        solver.sensing(U, sens)
        torch.cuda.synchronize()
        Xh, sh = X.cpu().numpy(), sens.cpu().numpy()
        ok = keep & np.isfinite(Xh).all(axis=1)
        out_X.append(Xh[ok])
        out_s.append(sh[ok])
    Xa = np.concatenate(out_X)
    sa = np.concatenate(out_s).astype(np.int32)
    return Dataset(X=Xa, y=(sa != 0).astype(np.float32), sensor=sa)


def interface_data(n_problems: int = 32, nx: int = 128, ny: int = 64, snapshots=(0, 5, 10, 20),
                   neg_ratio: int = 10, seed: int = SEED + 1, cfl: float = 0.3, device="cuda") -> Dataset:
    """Random two-fluid interface problems (multimode interfaces with random
    modes, amplitudes, density and gamma ratios, half of them hit by a shock
    of random Mach number), evolved with the global HR limiter; all flagged
    cells are kept with a seeded sample of `neg_ratio` unflagged cells per
    flagged one."""
    import torch

    rng = np.random.default_rng(seed)
    mesh = multimode_interface_mesh(nx, ny)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(SENSOR_THRESHOLDS)
    solver.set_indicator(Indicator(seed=0), 0.0)
    NE = mesh.sizes["N_E"]
    X = torch.zeros((NE, 64), dtype=torch.float32, device=device)
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    # cells next to the domain boundary are flagged by rule (by the sensor and
    # by the indicator alike) and carry no information: left out
    interior = ~(mesh.array("neighbors").reshape(-1, 4) < 0).any(axis=1)
    out_X, out_s = [], []
    for k in range(n_problems):
        U0 = multimode_interface_state(
            mesh, modes=int(rng.integers(1, 9)), mach=float(rng.uniform(1.05, 2.0)) if k % 2 else 0.0,
            seed=int(rng.integers(1 << 30)), rho_light=float(np.exp(rng.uniform(np.log(0.05), np.log(0.8)))),
            gamma_light=float(rng.uniform(1.1, 1.7)), amp_scale=float(rng.uniform(0.3, 3.0)))
        U = torch.from_numpy(U0).to(device)
        done = 0
        for snap in snapshots:
            if snap > done:
                if not _advance(solver, U, cfl * mesh.dx / 5, snap - done, LIMIT_HR):
                    break
                done = snap
            solver.features(U, X)
            solver.sensing(U, sens)
            torch.cuda.synchronize()
            Xh, sh = X.cpu().numpy(), sens.cpu().numpy()
            fin = np.isfinite(Xh).all(axis=1) & interior
            pos = np.nonzero(fin & (sh != 0))[0]
            neg = np.nonzero(fin & (sh == 0))[0]
            neg = rng.choice(neg, size=min(neg.size, max(neg_ratio * pos.size, 64)), replace=False)
            sel = np.concatenate([pos, neg])
            out_X.append(Xh[sel])
            out_s.append(sh[sel])
    Xa = np.concatenate(out_X)
    sa = np.concatenate(out_s).astype(np.int32)
    return Dataset(X=Xa, y=(sa != 0).astype(np.float32), sensor=sa)


def bubble_data(n_problems: int = 24, nx: int = 256, ny: int = 128, snapshots=(0, 10, 20, 40, 60, 80),
                neg_ratio: int = 10, seed: int = SEED + 3, cfl: float = 0.3, device="cuda") -> Dataset:
    """Random shocked bubbles on the shock-bubble box [0,2]x[0,1] (walls at
    y = 0 and 1): a planar shock of random Mach number in [1.05, 2] hits a
    circular bubble of random radius, centre, density (log-uniform in
    [0.05, 5]) and gamma ([1.1, 1.7]) in pressure equilibrium, evolved with
    the global HR limiter through the passage of the shock; flagged cells are
    kept with a seeded sample of unflagged ones, as in interface_data.  These
    teach the reflected and transmitted shocks a curved interface makes."""
    import torch

    rng = np.random.default_rng(seed)
    mesh = problems.shock_bubble_mesh(nx, ny, 2)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(SENSOR_THRESHOLDS)
    solver.set_indicator(Indicator(seed=0), 0.0)
    NE = mesh.sizes["N_E"]
    X = torch.zeros((NE, 64), dtype=torch.float32, device=device)
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    interior = ~(mesh.array("neighbors").reshape(-1, 4) < 0).any(axis=1)
    out_X, out_s = [], []
    for _ in range(n_problems):
        rb = float(rng.uniform(0.1, 0.3))
        U0 = problems.shock_bubble_state(
            mesh, mach=float(rng.uniform(1.05, 2.0)), xs=0.1, xb=float(rng.uniform(0.2 + rb, 1.2)),
            yb=float(rng.uniform(0.05 + rb, 0.95 - rb)), rb=rb,
            rho_b=float(np.exp(rng.uniform(np.log(0.05), np.log(5.0)))), gamma_b=float(rng.uniform(1.1, 1.7)))
        U = torch.from_numpy(U0).to(device)
        done = 0
        for snap in snapshots:
            if snap > done:
                if not _advance(solver, U, cfl * mesh.dx / 5, snap - done, LIMIT_HR):
                    break
                done = snap
            solver.features(U, X)
            solver.sensing(U, sens)
            torch.cuda.synchronize()
            Xh, sh = X.cpu().numpy(), sens.cpu().numpy()
            fin = np.isfinite(Xh).all(axis=1) & interior
            pos = np.nonzero(fin & (sh != 0))[0]
            neg = np.nonzero(fin & (sh == 0))[0]
            neg = rng.choice(neg, size=min(neg.size, max(neg_ratio * pos.size, 64)), replace=False)
            sel = np.concatenate([pos, neg])
            out_X.append(Xh[sel])
            out_s.append(sh[sel])
    Xa = np.concatenate(out_X)
    sa = np.concatenate(out_s).astype(np.int32)
    return Dataset(X=Xa, y=(sa != 0).astype(np.float32), sensor=sa)


def smooth_data(n_problems: int = 16, nx: int = 64, ny: int = 32, snapshots=(0, 5), amp_range=(0.005, 0.1),
                seed: int = SEED + 2, cfl: float = 0.3, device="cuda") -> Dataset:
    """Smooth random two-fluid flows on a periodic box (seeded waves of random
    amplitude in all fields and in gamma), evolved without limiting and
    labelled by the sensor: the smooth cells (most of them) teach the network
    that higher modes alone do not make a troubled cell."""
    import torch

    rng = np.random.default_rng(seed)
    mesh = Mesh(2, 2, nx, ny, 0.0, 2.0, 0.0, 1.0, bc=(PERIODIC,) * 4, nfields=5)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_sensor_thresholds(SENSOR_THRESHOLDS)
    solver.set_indicator(Indicator(seed=0), 0.0)
    NE = mesh.sizes["N_E"]
    X = torch.zeros((NE, 64), dtype=torch.float32, device=device)
    sens = torch.zeros(NE, dtype=torch.int32, device=device)
    out_X, out_s = [], []
    for _ in range(n_problems):
        amp = float(np.exp(rng.uniform(np.log(amp_range[0]), np.log(amp_range[1]))))
        U = torch.from_numpy(problems.smooth_random_state(mesh, MULTIFLUID, seed=int(rng.integers(1 << 30)),
                                                          amp=amp)).to(device)
        done = 0
        for snap in snapshots:
            if snap > done:
                if not _advance(solver, U, cfl * mesh.dx / 5, snap - done, 0):
                    break
                done = snap
            solver.features(U, X)
            solver.sensing(U, sens)
            torch.cuda.synchronize()
            Xh, sh = X.cpu().numpy(), sens.cpu().numpy()
            fin = np.isfinite(Xh).all(axis=1)
            out_X.append(Xh[fin])
            out_s.append(sh[fin])
    Xa = np.concatenate(out_X)
    sa = np.concatenate(out_s).astype(np.int32)
    return Dataset(X=Xa, y=(sa != 0).astype(np.float32), sensor=sa)


TRAINING_SETS = {
    "riemann": "4096 seeded random Riemann problems (64x64 mosaic of 8x8-cell blocks, rho and p log-uniform in "
               "[0.1, 10], u in [-2, 2], gamma in [1.1, 1.7]) and 16 smooth random two-fluid flows, labelled by "
               "the sensor",
    "interfaces": "the riemann set plus 32 random multimode interfaces (half shocked), labelled by the sensor",
    "full": "the interfaces set plus 24 random shocked bubbles, labelled by the sensor",
}


def training_set(kind: str = "riemann", device="cuda") -> Dataset:
    """Training data labelled by the sensor.  `riemann`: the seeded random
    Riemann mosaics and smooth random flows (the smooth flows teach that
    higher modes alone do not make a troubled cell); neither the shock-bubble
    nor a multimode interface is in it, so both are out-of-sample
    evaluations.  `interfaces` adds random multimode interfaces, `full` also
    random shocked bubbles."""
    if kind not in TRAINING_SETS:
        raise ValueError(f"training_set: unknown kind {kind!r}")
    parts = [generate(device=device), smooth_data(device=device)]
    if kind in ("interfaces", "full"):
        parts.append(interface_data(device=device))
    if kind == "full":
        parts.append(bubble_data(device=device))
    return merge(*parts)


def merge(*parts: Dataset) -> Dataset:
    return Dataset(X=np.concatenate([p.X for p in parts]), y=np.concatenate([p.y for p in parts]),
                   sensor=np.concatenate([p.sensor for p in parts]))


def train(data: Dataset, seed: int = SEED, epochs: int = 30, batch: int = 1024, lr: float = 2e-3, device="cuda"):
    """Train a fresh indicator on the GPU; returns (network, epoch losses,
    threshold).  The threshold is set below the lowest logit of any training
    cell the sensor flags as a shock, so that no shock cell is missed."""
    import torch

    net = Indicator(seed=seed & 0xFFFFFFFF)
    Xd = torch.from_numpy(data.X).to(device)
    yd = torch.from_numpy(data.y).to(device)
    pos = max(float(data.y.sum()), 1.0)
    pos_weight = (data.y.size - pos) / pos
    losses = net.train(Xd, yd, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    z = logits(net, data.X, device)
    shock = data.sensor == 2
    thr = float(np.min(z[shock]) - 1e-3) if shock.any() else 0.0
    return net, losses, min(thr, 0.0)


def logits(net: Indicator, X: np.ndarray, device="cuda") -> np.ndarray:
    import torch

    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(device)
    z = torch.zeros(X.shape[0], dtype=torch.float32, device=device)
    net.forward(Xd, z)
    torch.cuda.synchronize()
    return z.cpu().numpy()


def confusion(flags: np.ndarray, sensor: np.ndarray) -> dict:
    """The plain sensor judges the indicator: missed = flagged by the sensor
    only, spurious = flagged by the indicator only, per sensor kind (cells
    next to a domain boundary are flagged by rule by both)."""
    f = np.asarray(flags) != 0
    s = np.asarray(sensor)
    return {
        "cells": int(s.size), "sensor_flagged": int((s != 0).sum()), "indicator_flagged": int(f.sum()),
        "missed": int(((s != 0) & ~f).sum()), "spurious": int((f & (s == 0)).sum()),
        "missed_shock": int(((s == 2) & ~f).sum()), "shock_cells": int((s == 2).sum()),
    }


def multimode_interface_mesh(nx: int = 2048, ny: int = 1024) -> Mesh:
    return Mesh(2, 2, nx, ny, 0.0, 2.0, 0.0, 1.0, bc=(FARFIELD, FARFIELD, FARFIELD, FARFIELD), nfields=5)


def multimode_interface_state(mesh: Mesh, modes: int = 8, mach: float = 1.22, seed: int = SEED,
                              rho_light: float = 0.138, gamma_light: float = 1.667,
                              amp_scale: float = 1.0) -> np.ndarray:
    """A two-fluid interface x = 0.6 + sum_k a_k sin(2 pi k y + phi_k) between
    air (rho=1, gamma=1.4, left) and a light gas (rho=0.138, gamma=1.667,
    right) in pressure equilibrium, hit by a Mach `mach` shock at x = 0.3
    (no shock when mach <= 1)."""
    rng = np.random.default_rng(seed)
    x, y = problems.node_xy(mesh)
    xe, ye = x[:, :1] * np.ones_like(x), y[:, :1] * np.ones_like(y)
    amp = amp_scale * rng.uniform(0.005, 0.02, modes) / np.arange(1, modes + 1)
    ph = rng.uniform(0.0, 2 * math.pi, modes)
    xi = 0.6 + sum(amp[k] * np.sin(2 * math.pi * (k + 1) * ye + ph[k]) for k in range(modes))
    g = 1.4
    mach = max(mach, 1.0)
    M2 = mach * mach
    p2 = 1.0 + 2 * g / (g + 1) * (M2 - 1.0)
    rho2 = (g + 1) * M2 / ((g - 1) * M2 + 2.0)
    u2 = mach * math.sqrt(g) * (1.0 - 1.0 / rho2)
    post = xe < (0.3 if mach > 1.0 else -1.0)
    light = xe > xi
    rho = np.where(post, rho2, np.where(light, rho_light, 1.0))
    u = np.where(post, u2, 0.0)
    p = np.where(post, p2, 1.0)
    gam = np.where(light & ~post, gamma_light, 1.4)
    E = p / (gam - 1.0) + 0.5 * rho * u * u
    f = np.stack([rho, rho * u, np.zeros_like(rho), E, 1.0 / (gam - 1.0)], axis=1)
    return np.ascontiguousarray(f.reshape(-1))


def _profile_pass(nx: int = 1024, ny: int = 512, passes: int = 4) -> None:
    """The flagging pass alone on the shock-bubble state, for profilers: a
    seeded untrained network (the pass costs the same whatever the weights)
    and no training, so that a counter run sees a handful of dispatches."""
    import torch

    mesh = problems.shock_bubble_mesh(nx, ny, 2)
    solver = Solver(Config(2, 2, MULTIFLUID, ROE), mesh)
    solver.set_indicator(Indicator(seed=0), 0.0)
    U = torch.from_numpy(problems.shock_bubble_state(mesh)).to("cuda")
    flags = torch.zeros(mesh.sizes["N_E"], dtype=torch.int32, device="cuda")
    for _ in range(passes):
        solver.indicator_flags(U, flags)
    torch.cuda.synchronize()
    print(f"indicator pass x{passes} on {nx}x{ny}: {int(flags.sum().item())} cells flagged", flush=True)


if __name__ == "__main__":
    _profile_pass()
