#!/usr/bin/env python
"""DOF-updates/s of the DG RK4 step on one MI355X.

Default workload (bubble): the 2D two-fluid shock-bubble (Mach 1.22 shock in
gamma=1.4 gas hitting a circular bubble of radius 0.25, rho=0.138,
gamma=1.667) on [0,2]x[0,1] with 1024x512 quads, tensor Lagrange p=2, 5
fields, Roe flux, RK4 with a fixed Dt from CFL 0.4 at t=0.  The default run
times the reduced-precision form (FP32 map/integrate contractions on the
matrix cores, FP64 fluxes, RK and limiter; HR applied to the cells the
learned indicator flags, the indicator trained on the GPU before the timed
region) and, in the same run and on the same initial state, the plain FP64
form (FP64 contractions, global HR limiter): `value` is the reduced form's
throughput, `plain_fp64_form` the plain one's and their ratio.  One
DOF-update = one conserved coefficient advanced one full RK4 step (four
right-hand sides).  Synthetic input (the initial condition, interpolated at
the nodes).

Other workloads (--config): `vortex`, the 2D single-fluid isentropic vortex on
a periodic [0,10]^2 mesh of 512x512 quads, p=2, 4 fields, Roe, no limiter,
fixed Dt from CFL 0.3; `sod`, the 1D single-fluid Sod tube on 1000 elements,
p=2, Rusanov, HR limiter, CFL 0.4 steps (Dt from the state every step);
`ensemble`, 256 independent 1D two-fluid Riemann problems (element counts
uniform in [200, 4000], orders uniform in {1, 2, 3}, seeded states), Roe, HR
limiter, each with its own CFL 0.4 step, advanced by one graph launch of K
steps.  `--precision` and `--limiter` select another form (on `bubble`:
`--precision fp64 --limiter hr` is the plain form alone).

    python bench.py [--steps K] [--warmup W] [--config bubble|vortex|sod|ensemble] [--no-cpu-baseline]

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
PRECISION_MODE = {"fp64": 0, "fp32mfma": 1, "splitbf16": 2}  # dgx_solver_set_precision


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", choices=["bubble", "vortex", "sod", "ensemble"], default="bubble")
    ap.add_argument("--n", type=int, default=0, help="vortex: elements per direction (default 512)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["auto", "fp64", "fp32mfma", "splitbf16"], default="auto",
                    help="tile-kernel volume contractions: FP64, FP32 on MFMA, or split-BF16 products with FP32 "
                         "accumulation (2D p=2; the rest stays FP64); auto: FP32 on the bubble, FP64 elsewhere")
    ap.add_argument("--path", choices=["auto", "general", "tiled"], default="auto",
                    help="kernel path: fused tile kernel (auto on this mesh) or face + element kernels")
    ap.add_argument("--limiter", choices=["auto", "none", "hr", "hri", "learned"], default="auto",
                    help="limiter between stages: auto (learned on the bubble, else the workload's own), none, "
                         "global HR, sensor-gated HR, or HR gated by the learned indicator (trained on the GPU before "
                         "the timed run; 2D two-fluid p=2)")
    ap.add_argument("--training", choices=["riemann", "interfaces", "full"], default="riemann",
                    help="indicator training data: seeded random Riemann mosaics and smooth flows (riemann), plus "
                         "random multimode interfaces (interfaces), plus random shocked bubbles (full)")
    ap.add_argument("--cpu-sample-steps", type=int, default=0, help="CPU sample steps (0: workload default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB buffer 5 times (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.precision == "auto":
        args.precision = "fp32mfma" if args.config == "bubble" else "fp64"
    if args.limiter == "auto" and args.config == "bubble":
        args.limiter = "learned"

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs (--gpus N > 1) are not supported",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from dg_amd import PATH_AUTO, PATH_GENERAL, PATH_TILED, Config, Solver

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

    if args.config == "ensemble":
        return bench_ensemble(args, device)
    w = workload(args)
    mesh, order = w["mesh"], w["order"]
    cfg = Config(w["dim"], order, w["physics"], w["flux"])
    solver = Solver(cfg, mesh)
    solver.set_path({"auto": PATH_AUTO, "general": PATH_GENERAL, "tiled": PATH_TILED}[args.path])
    tiled = solver.path == PATH_TILED
    if args.precision != "fp64":
        solver.set_precision(PRECISION_MODE[args.precision])
    s = mesh.sizes
    N_E, N_s, NF = s["N_E"], s["N_s"], cfg.nfields
    n_dof = N_E * N_s * NF

    U0 = w["U0"]
    U = torch.from_numpy(U0).to(device)
    cfl = w["cfl"] * mesh.dx / (2 * order + 1)
    dt_dev = torch.zeros(1, dtype=torch.float64, device=device)
    solver.dt_from_cfl(U, cfl, dt_dev, stream=stream)
    torch.cuda.synchronize()
    dt = float(dt_dev.item())
    lim = {"auto": w["limiter"], "none": 0, "hr": 1, "hri": 3, "learned": 8}[args.limiter]
    extra = {}
    net = None
    if lim in (3, 8):
        from dg_amd import indicator
        solver.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
    if lim == 8:  # train the indicator on the GPU (own forward/backward kernels), outside the timed region
        t_tr = time.perf_counter()
        data = indicator.training_set(args.training)
        net, losses, thr = indicator.train(data, epochs=40)
        solver.set_indicator(net, thr)
        extra["indicator"] = {"training_s": round(time.perf_counter() - t_tr, 2), "samples": int(data.y.size),
                              "final_loss": float(losses[-1]), "threshold": thr,
                              "training_data": indicator.TRAINING_SETS[args.training]}

    def make_advance(sv, Ux, lim_x):
        def adv(k):
            if w["cfl_steps"]:
                sv.rk4_integrate(Ux, 1, cfl, nsteps=k, tf=1e30, limiter=lim_x, stream=stream)
            else:
                sv.rk4_integrate(Ux, 0, dt, nsteps=k, limiter=lim_x, stream=stream)
        return adv

    advance = make_advance(solver, U, lim)

    # warmup, then exactly K timed steps bracketed by device synchronisation
    advance(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    advance(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_dof * args.steps / elapsed

    if lim in (3, 8):  # the plain sensor judges the flags of the final state
        from dg_amd import indicator
        sens = torch.zeros(N_E, dtype=torch.int32, device=device)
        solver.sensing(U, sens)
        if lim == 8:
            fl = torch.zeros_like(sens)
            solver.indicator_flags(U, fl)
            torch.cuda.synchronize()
            extra["indicator"]["flags_vs_sensor_final_state"] = indicator.confusion(fl.cpu().numpy(),
                                                                                     sens.cpu().numpy())
        else:
            torch.cuda.synchronize()
            extra["sensor_flagged_final_state"] = int((sens.cpu().numpy() != 0).sum())
    if args.precision != "fp64" or lim in (3, 8):
        # the plain FP64 GPU form (global HR) on the same workload and steps, timed here too
        sp = Solver(cfg, mesh)
        Up = torch.from_numpy(U0).to(device)
        adv_p = make_advance(sp, Up, w["limiter"])
        adv_p(args.warmup)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        adv_p(args.steps)
        torch.cuda.synchronize()
        ms_plain = 1e3 * (time.perf_counter() - tp) / args.steps
        torch.cuda.synchronize()
        ap_ = Up.cpu().numpy().reshape(N_E, NF, N_s)
        af_ = U.cpu().numpy().reshape(N_E, NF, N_s)
        extra["plain_fp64_form"] = {
            "limiter": {0: "none", 1: "global HR"}.get(w["limiter"], w["limiter"]), "contractions": "FP64",
            "ms_per_step": ms_plain, "value": n_dof * args.steps / (ms_plain * 1e-3 * args.steps),
            "speedup_of_this_form": ms_plain / ms_per_step, "target_speedup": 2.0,
            # the two forms' states after the same warmup + K steps (limiter and precision differ)
            "rel_l2_this_form_vs_plain": float(np.linalg.norm(af_ - ap_) / np.linalg.norm(ap_))}
    if args.precision != "fp64":
        # precision error: the same limiter in FP64 from the same state, warmup + K steps
        sf = Solver(cfg, mesh)
        if lim in (3, 8):
            sf.set_sensor_thresholds(indicator.SENSOR_THRESHOLDS)
        if lim == 8:
            sf.set_indicator(net, extra["indicator"]["threshold"])
        Uf = torch.from_numpy(U0).to(device)
        make_advance(sf, Uf, lim)(args.warmup + args.steps)
        torch.cuda.synchronize()
        a64 = Uf.cpu().numpy().reshape(N_E, NF, N_s)
        a32 = U.cpu().numpy().reshape(N_E, NF, N_s)
        nrm = float(np.linalg.norm(a64))
        extra["rel_l2_vs_fp64_same_limiter"] = {
            "state": float(np.linalg.norm(a32 - a64) / nrm),
            "fields_rel_to_state_norm": [float(np.linalg.norm(a32[:, f] - a64[:, f]) / nrm) for f in range(NF)]}

    if args.config == "vortex":  # accuracy against the exact translated vortex after warmup + K steps
        from dg_amd import problems
        t_final = (args.warmup + args.steps) * dt
        Uh = U.cpu().numpy().reshape(N_E, NF, N_s)
        x, y = problems.node_xy(mesh)
        rho_ex, _, _, _ = problems.vortex_fields(x, y, t=t_final)
        extra["l2_rel_err_rho_vs_exact"] = float(np.linalg.norm(Uh[:, 0] - rho_ex) / np.linalg.norm(rho_ex))

    # per-kernel timing with HIP events on the launch stream (separate run)
    prof_steps = max(1, min(args.steps, 20))
    names = ("stage", "elem", "face", "limit", "dt")
    solver.profile(True)
    for k in names:
        solver.profile_read(k)
    advance(prof_steps)
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
    rhs_kernels = {k: v for k, v in avg.items() if k in ("k_stage", "k_elem", "k_face")}
    dominant = max(rhs_kernels, key=rhs_kernels.get)
    dom_avg = avg[dominant]
    dom_bytes = face_bytes if dominant == "k_face" else stage_bytes
    achieved = dom_bytes / (dom_avg * 1e-3) / 1e9
    # whole-step view: minimum RK4 state traffic per step, in state vectors:
    # 16 for the fused stage flow, plus 2 per limiter sweep (read + write) for
    # each of the 4 limiter applications (1 sweep in 1D, 2 in 2D)
    # (flag-gated limiters: one state read per application for the sensor or
    # indicator, the flagged cells' sweeps not counted)
    c = 16 + (0 if not lim else 4 * 2 * w["dim"] if lim == 1 else 4)
    step_gbps = c * 8.0 * n_dof / (ms_per_step * 1e-3) / 1e9

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
        "dtype": {"fp64": "f64", "fp32mfma": "f64+f32 (mfma contractions)",
                  "splitbf16": "f64+split-bf16 (mfma contractions, f32 accumulation)"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": w["name"],
            "elements": N_E, "dofs": n_dof, "dt": dt, "parallelism": "single GPU",
            "kernel_path": "tiled (one fused kernel per RK stage)" if tiled else "face + element kernels",
            "limiter": {0: "none", 1: "global HR", 3: "sensor-gated HR", 8: "HR gated by the learned indicator"}[lim],
            "contractions": ("FP32 MFMA (FP64 fluxes, RK and limiter)" if args.precision == "fp32mfma" else
                             "split-BF16 MFMA, FP32 accumulation (FP64 fluxes, RK and limiter)"
                             if args.precision == "splitbf16" else
                             "FP64 MFMA element phases" if tiled and w["physics"] == 0 and w["order"] == 2 and w["dim"] == 2
                             else "FP64"),
        },
        "roofline": {
            "bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": (measured_traffic(args.config if args.precision == "fp64" else f"{args.config}-{args.precision}",
                                         dominant)
                        if args.path != "general" and args.n in (0, 512) else None),
            "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_avg,
            "kernels_ms": avg,
        },
        "step_roofline": {"c": c, "achieved": step_gbps, "frac": step_gbps / HBM_PEAK_GBPS, "unit": "GB/s"},
    }
    out.update(extra)

    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["l2_rel_err_vs_oracle"] = cpu_baseline(args, device, w)

    print(json.dumps(out))
    return 0


def bench_ensemble(args, device):
    """The batched ensemble: autotune, W warmup steps, then K steps timed as
    one graph launch.  DOF-updates count the problems that completed every
    step (status 0); strong random Riemann problems can lose positivity and
    stop, as they would in the reference."""
    import numpy as np
    import torch

    import oracle
    from dg_amd import FARFIELD, LIMIT_HR, MULTIFLUID, ROE, Batch, Config, Mesh, problems

    n, cfl = 256, 0.4
    ne, p = problems.ensemble_spec(n, seed=20261202)
    meshes = [Mesh(1, p[k], ne[k], 1, -1.0, 1.0, bc=(FARFIELD, FARFIELD), nfields=4) for k in range(n)]
    states = [problems.two_fluid_riemann_1d(meshes[k], 20261202 + k) for k in range(n)]
    batch = Batch(Config(1, 1, MULTIFLUID, ROE), ne, p)
    U = torch.from_numpy(batch.pack(states)).to(device)
    T = torch.zeros(n, dtype=torch.float64, device=device)
    status = torch.zeros(n, dtype=torch.int32, device=device)
    batch.autotune(U, cfl)
    batch.rk4(U, cfl, args.warmup, T, status, limit_initial=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.rk4(U, cfl, args.steps, T, status, limit_initial=False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = status.cpu().numpy()
    dofs = np.array([batch.offsets[k + 1] - batch.offsets[k] for k in range(n)])
    ok = st == 0
    value = float(dofs[ok].sum()) * args.steps / elapsed
    out = {
        "metric": "DOF-updates/s (1 GPU) at fixed L2 err vs FP64 oracle; % HBM/MFMA roofline",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "256 independent 1D two-fluid Riemann problems (N_E uniform in [200,4000], p uniform "
                               "in {1,2,3}, rho,p log-uniform [0.1,10], u in [-2,2], gamma in [1.1,1.7] per side), "
                               "Roe, HR limiter, per-problem Dt from CFL 0.4, one graph launch of K steps",
                   "problems": n, "dofs": int(dofs.sum()), "completed": int(ok.sum()),
                   "block_sizes": batch.block_sizes, "parallelism": "single GPU"},
        "step_roofline": {"c": 16 + 8, "achieved": 24 * 8.0 * float(dofs[ok].sum()) / (elapsed / args.steps) / 1e9,
                          "unit": "GB/s"},
    }
    out["step_roofline"]["frac"] = out["step_roofline"]["achieved"] / HBM_PEAK_GBPS
    if not args.no_cpu_baseline:
        # oracle on a bounded sample: the first 16 problems that complete, 20 steps each
        Uh = batch.unpack(U.cpu().numpy())
        t_cpu, dof_steps, errs = 0.0, 0, []
        for k in [k for k in range(n) if ok[k]][:16]:
            prob = oracle.Problem(meshes[k].tables(), MULTIFLUID, ROE)
            c = cfl * meshes[k].dx / (2 * p[k] + 1)
            t1 = time.perf_counter()
            _, steps, _, _ = prob.rk4(states[k], 1, c, nsteps=20, tf=1e30, limiter=LIMIT_HR)
            t_cpu += time.perf_counter() - t1
            dof_steps += states[k].size * max(steps, 0)
        out["cpu_baseline"] = {"value": dof_steps / t_cpu, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
                               "sample": f"16 of the problems, 20 RK4 steps each, FP64 oracle, {t_cpu:.1f} s",
                               **host_cpu()}
    print(json.dumps(out))
    return 0


def measured_traffic(config: str, kernel: str):
    """HBM bytes per dispatch of `kernel` on `config` from the committed
    rocprofv3 --pmc table (profiles/traffic.txt: FETCH_SIZE and WRITE_SIZE in
    KiB with their copy-calibrated scales), or None when it has no row."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    scale, rows = {}, {}
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] == "scale":
            scale[f[1]] = float(f[2])
        else:
            rows[(f[0], f[1])] = (float(f[2]), float(f[3]))
    if (config, kernel) not in rows or len(scale) != 2:
        return None
    fetch, write = rows[(config, kernel)]
    return 1024.0 * (scale["FETCH_SIZE"] * fetch + scale["WRITE_SIZE"] * write)


def workload(args):
    """Mesh, configuration, initial state and stepping of the selected workload."""
    from dg_amd import MULTIFLUID, ROE, RUSANOV, SINGLEFLUID, problems

    if args.config == "vortex":
        n = args.n or 512
        mesh = problems.vortex_mesh(n, 2)
        return dict(name=f"2D isentropic vortex, periodic [0,10]^2, {n}x{n} quads, p=2 Lagrange, 4 fields, Roe, "
                         f"no limiter, RK4 fixed Dt from CFL 0.3", mesh=mesh, dim=2, order=2, physics=SINGLEFLUID,
                    flux=ROE, U0=problems.vortex_state(mesh), cfl=0.3, limiter=0, cfl_steps=False,
                    sample=lambda: problems.vortex_mesh(n, 2), sample_state=problems.vortex_state, sample_steps=4)
    if args.config == "sod":
        mesh = problems.sod_mesh(1000, 2, SINGLEFLUID)
        return dict(name="1D single-fluid Sod tube, [-1,1], 1000 elements, p=2 Lagrange, 3 fields, Rusanov, HR "
                         "limiter, RK4 with Dt from CFL 0.4 every step", mesh=mesh, dim=1, order=2,
                    physics=SINGLEFLUID, flux=RUSANOV, U0=problems.sod_state(mesh, SINGLEFLUID), cfl=0.4, limiter=1,
                    cfl_steps=True, sample=lambda: problems.sod_mesh(1000, 2, SINGLEFLUID),
                    sample_state=lambda m: problems.sod_state(m, SINGLEFLUID), sample_steps=200)
    mesh = problems.shock_bubble_mesh(1024, 512, 2)
    return dict(name="2D two-fluid shock-bubble (Mach 1.22, bubble r=0.25 rho=0.138 gamma=1.667), [0,2]x[0,1], "
                     "1024x512 quads, p=2 Lagrange, 5 fields, Roe, HR limiter, RK4 fixed Dt from CFL 0.4 at t=0",
                mesh=mesh, dim=2, order=2, physics=MULTIFLUID, flux=ROE, U0=problems.shock_bubble_state(mesh),
                cfl=0.4, limiter=1, cfl_steps=False, sample=lambda: problems.shock_bubble_mesh(1024, 512, 2),
                sample_state=problems.shock_bubble_state, sample_steps=2)


def cpu_baseline(args, device, w):
    """FP64 CPU oracle (one core) on a bounded sample of the same workload
    (a few RK4 steps); the GPU engine runs the same sample and the largest
    per-field relative L2 difference is reported."""
    import numpy as np
    import torch

    import oracle
    from dg_amd import Config, Solver

    mesh = w["sample"]()
    prob = oracle.Problem(mesh.tables(), w["physics"], w["flux"])
    U0 = w["sample_state"](mesh)
    order = w["order"]
    cfl = w["cfl"] * mesh.dx / (2 * order + 1)
    nst = args.cpu_sample_steps if args.cpu_sample_steps > 0 else w["sample_steps"]
    dt = prob.dt_from_cfl(U0, cfl)
    t0 = time.perf_counter()
    if w["cfl_steps"]:
        U_ref, steps, _, _ = prob.rk4(U0, 1, cfl, nsteps=nst, tf=1e30, limiter=w["limiter"])
    else:
        U_ref, steps, _, _ = prob.rk4(U0, 0, dt, nsteps=nst, limiter=w["limiter"])
    el = time.perf_counter() - t0
    solver = Solver(Config(w["dim"], order, w["physics"], w["flux"]), mesh)
    if args.precision != "fp64":
        solver.set_precision(PRECISION_MODE[args.precision])
    Ud = torch.from_numpy(U0).to(device)
    if w["cfl_steps"]:
        solver.rk4_integrate(Ud, 1, cfl, nsteps=nst, tf=1e30, limiter=w["limiter"])
    else:
        solver.rk4_integrate(Ud, 0, dt, nsteps=nst, limiter=w["limiter"])
    torch.cuda.synchronize()
    Ug = Ud.cpu().numpy()
    nf, ns = prob.nf, mesh.sizes["N_s"]
    A, B = Ug.reshape(-1, nf, ns), U_ref.reshape(-1, nf, ns)
    # relative L2 error of the whole state, and of each field relative to the
    # whole state's norm (a field that is zero in exact arithmetic, such as
    # rho*v in a planar problem, has no meaningful error relative to itself)
    nrm = float(np.linalg.norm(B))
    err = {"state": float(np.linalg.norm(A - B)) / nrm,
           "fields": [float(np.linalg.norm(A[:, f] - B[:, f])) / nrm for f in range(nf)]}
    cpu = {
        "value": U0.size * steps / el, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
        "sample": f"{steps} RK4 step(s) of the same workload ({U0.size} DOFs), FP64 oracle "
                  f"(oracle/dg_oracle.cc, g++ -O2), {el:.1f} s",
        **host_cpu(),
    }
    return cpu, err


def host_cpu() -> dict:
    """The host the CPU baseline ran on: logical CPUs visible to this process
    and the CPU model string."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        visible = os.cpu_count()
    return {"host_nproc": visible, "host_cpu_model": model}


if __name__ == "__main__":
    sys.exit(main())
