#!/usr/bin/env python
"""DOF-updates/s of the FP64 DG RK4 step on one MI355X.

Workload: 2D single-fluid isentropic vortex on a periodic [0,10]^2 mesh of
512x512 quads, tensor Lagrange p=2 (9 nodes x 4 fields per element), Roe
flux, no limiting, classic RK4 with a fixed Dt from CFL 0.3 (deck CFL x
Dx/(2p+1) / max(|u|+a)).  Synthetic input (the exact vortex, interpolated at
the nodes).  One DOF-update = one conserved coefficient advanced one full RK4
step (four right-hand sides).

    python bench.py [--steps K] [--warmup W] [--n 512] [--no-cpu-baseline]

Prints one JSON line.  `value` is the timed region's throughput with the
state already resident in HBM; `roofline` is the dominant kernel's achieved
HBM bandwidth measured with HIP events on the launch stream; `cpu_baseline`
is the FP64 CPU oracle timed on this host (one core) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBPS = 8000.0  # MI355X published HBM3E bandwidth


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="elements per direction")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--path", choices=["auto", "general", "tiled"], default="auto",
                    help="kernel path: fused tile kernel (auto on this mesh) or face + element kernels")
    ap.add_argument("--cpu-sample-n", type=int, default=0, help="CPU sample mesh size (0: same as --n)")
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB buffer 5 times (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs (--gpus N > 1) are not part of this build "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from dg_amd import PATH_AUTO, PATH_GENERAL, PATH_TILED, ROE, SINGLEFLUID, Config, Solver, problems

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    device = torch.device("cuda:0")
    stream = torch.cuda.current_stream(device)

    if args.copy_calibration:
        nbytes = 256 << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=device)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "copy_calibration_bytes_written": nbytes,
                          "dispatches": 5}))
        return 0

    order, n = 2, args.n
    mesh = problems.vortex_mesh(n, order)
    cfg = Config(2, order, SINGLEFLUID, ROE)
    solver = Solver(cfg, mesh)
    solver.set_path({"auto": PATH_AUTO, "general": PATH_GENERAL, "tiled": PATH_TILED}[args.path])
    tiled = solver.path == PATH_TILED
    s = mesh.sizes
    N_E, N_s, NF = s["N_E"], s["N_s"], cfg.nfields
    n_dof = N_E * N_s * NF

    U0 = problems.vortex_state(mesh)
    U = torch.from_numpy(U0).to(device)
    cfl = 0.3 * mesh.dx / (2 * order + 1)
    dt_dev = torch.zeros(1, dtype=torch.float64, device=device)
    solver.dt_from_cfl(U, cfl, dt_dev, stream=stream)
    torch.cuda.synchronize()
    dt = float(dt_dev.item())

    # warmup, then exactly K timed steps bracketed by device synchronisation
    solver.rk4_integrate(U, 0, dt, nsteps=args.warmup, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.rk4_integrate(U, 0, dt, nsteps=args.steps, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_dof * args.steps / elapsed

    # accuracy against the exact translated vortex after warmup + K steps
    t_final = (args.warmup + args.steps) * dt
    Uh = U.cpu().numpy().reshape(N_E, NF, N_s)
    x, y = problems.node_xy(mesh)
    rho_ex, _, _, _ = problems.vortex_fields(x, y, t=t_final)
    l2_exact = float(np.linalg.norm(Uh[:, 0] - rho_ex) / np.linalg.norm(rho_ex))

    # per-kernel timing with HIP events on the launch stream (separate run)
    prof_steps = max(1, min(args.steps, 20))
    names = ("stage", "elem", "face")
    solver.profile(True)
    for k in names:
        solver.profile_read(k)
    solver.rk4_integrate(U, 0, dt, nsteps=prof_steps, stream=stream)
    torch.cuda.synchronize()
    avg = {}
    for k in names:
        ms, cnt = solver.profile_read(k)
        if cnt:
            avg["k_" + k] = ms / cnt
    solver.profile(False)
    # Algorithmic bytes of one stage launch (k_stage, or k_elem on the general
    # path): the state vectors of the reference RK data flow with DU fused
    # away, averaged over the four stages: stage 0 R(U) W(Uacc) W(U*) = 3,
    # stages 1-2 R(U*) R(U) R(Uacc) W(Uacc) W(U*) = 5, stage 3 R(U*) R(Uacc)
    # W(U) = 3 -> 4 state vectors of n_dof doubles per launch.  Halo traces
    # and geometry are not counted.
    stage_bytes = 4.0 * 8.0 * n_dof
    # Face kernel: one read of the face traces (<= one state vector) and its
    # per-face nodal contributions (2 sides x N_F x M_s per face).
    face_bytes = 8.0 * (n_dof + 2 * NF * s["M_s"] * s["M_T"])
    dominant = max(avg, key=avg.get)
    dom_avg = avg[dominant]
    dom_bytes = face_bytes if dominant == "k_face" else stage_bytes
    achieved = dom_bytes / (dom_avg * 1e-3) / 1e9
    # whole-step view: minimum RK4 state traffic, c = 16 state vectors per step
    step_gbps = 16.0 * 8.0 * n_dof / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "DOF-updates/s (1 GPU) at fixed L2 err vs FP64 oracle; % HBM/MFMA roofline",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"2D isentropic vortex, periodic [0,10]^2, {n}x{n} quads, p=2 Lagrange, 4 fields, Roe, "
                        f"no limiter, RK4 fixed Dt from CFL 0.3",
            "elements": N_E, "dofs": n_dof, "dt": dt, "parallelism": "single GPU",
            "kernel_path": "tiled (one fused kernel per RK stage)" if tiled else "face + element kernels",
        },
        "roofline": {
            "bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
            "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_avg,
            "kernels_ms": avg,
        },
        "step_roofline": {"c": 16, "achieved": step_gbps, "frac": step_gbps / HBM_PEAK_GBPS, "unit": "GB/s"},
        "l2_rel_err_rho_vs_exact": l2_exact,
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["l2_rel_err_vs_oracle"] = cpu_baseline(args, device)

    print(json.dumps(out))
    return 0


def cpu_baseline(args, device):
    """FP64 CPU oracle (one core) on a bounded sample of the same workload:
    `cpu-sample-steps` RK4 steps of the vortex at cpu-sample-n^2; the GPU
    engine runs the same sample and the largest per-field relative L2
    difference is reported."""
    import numpy as np
    import torch

    import oracle
    from dg_amd import PATH_AUTO, PATH_GENERAL, PATH_TILED, ROE, SINGLEFLUID, Config, Solver, problems

    n = args.cpu_sample_n if args.cpu_sample_n else args.n
    mesh = problems.vortex_mesh(n, 2)
    prob = oracle.Problem(mesh.tables(), SINGLEFLUID, ROE)
    U0 = problems.vortex_state(mesh)
    dt = prob.dt_from_cfl(U0, 0.3 * mesh.dx / 5)
    t0 = time.perf_counter()
    U_ref, steps, _, _ = prob.rk4(U0, 0, dt, nsteps=args.cpu_sample_steps)
    el = time.perf_counter() - t0
    solver = Solver(Config(2, 2, SINGLEFLUID, ROE), mesh)
    Ud = torch.from_numpy(U0).to(device)
    solver.rk4_integrate(Ud, 0, dt, nsteps=args.cpu_sample_steps)
    torch.cuda.synchronize()
    Ug = Ud.cpu().numpy()
    nf, ns = 4, mesh.sizes["N_s"]
    A, B = Ug.reshape(-1, nf, ns), U_ref.reshape(-1, nf, ns)
    err = max(float(np.linalg.norm(A[:, f] - B[:, f]) / np.linalg.norm(B[:, f])) for f in range(nf))
    cpu = {
        "value": U0.size * steps / el, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
        "sample": f"{steps} RK4 step(s) of the same vortex workload at {n}x{n} ({U0.size} DOFs), "
                  f"FP64 oracle (oracle/dg_oracle.cc, g++ -O2), {el:.1f} s",
    }
    return cpu, err


if __name__ == "__main__":
    sys.exit(main())
