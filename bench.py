#!/usr/bin/env python3
"""Throughput benchmark of the 3D WENO-4 / CWENO4 right-hand side + SSP-RK3
stepper on one MI355X.

Metric: RK-stage cell-updates/s (cells x stages / wall time), one
cell-update = one RHS evaluation of one cell fused with its stage
combination; the stepper is the 4-stage SSP RK3 (TSSSPRKS3).

Headline workload: 3D linearised acoustics (p', u, v, w), rho0 = c0 = 1, on a
256^3 periodic grid of the unit cube, FP64, CFL 0.4, initial field a sum of 4
seeded Gaussian pressure pulses (amplitudes U[1e-4,1e-3], sigma
U[0.02,0.08]), zero velocity; shipped linear weights gamma = (1, 0).  The
state is resident in HBM before timing starts.

Also reported ("variants"): the nonlinear CWENO4 (gamma = 0.85/0.85) FP64
kernel on the same workload, and the FP32 compensated path at 384^3 on the
band-limited noise + pulse field, with its error against an FP64 run.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
                       [--variants all|none] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK-stage cell-updates/s, 3D WENO4: FP64 plain vs FP32+learned; % HBM BW"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
STAGES = 4
# Algorithmic HBM bytes per cell-update (compulsory state traffic of the
# low-storage 4-stage SSP RK3: 5 state reads + 4 state writes per step):
BYTES_FP64_AC4 = 9 * 4 * 8 / 4.0        # 72 B


def f32c_bytes_per_update():
    # per step, 4 components x 4 B = 16 B per cell and array:
    # stage 1: read HI, write D0; stage 2: read HI, D0, write D1;
    # stage 3: read HI, D1, write D0; stage 4: read HI, D0, LO, write HI', LO'
    reads = 1 + 2 + 2 + 3
    writes = 1 + 1 + 1 + 2
    return (reads + writes) * 16 / 4.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--variants", default="all", choices=["all", "none"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain device copy of the state (counter calibration)")
    return ap.parse_args()


def traffic_from_profiles(n):
    """HBM bytes per launch of the headline kernel from the committed PMC summary."""
    path = os.path.join(ROOT, "profiles", f"k_linear_fp64_ac4_{n}_pmc.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1].strip())
                except ValueError:
                    return None
    return None


def time_steps(torch, step, steps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        step()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1e3, wall


def cpu_baseline(acx, n_sample=96, max_steps=8, budget_s=10.0):
    """The FP64 oracle (reference loop structure: ghosted copy, per-Gauss-point
    scratch, per-Gauss-point LLF) timed on this host on a 96^3 sample of the
    same workload."""
    from oracle import oracle as orc
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    orc.lib().orc_set_threads(threads)
    grid = acx.Grid(n=(n_sample,) * 3, h=1.0 / n_sample, system=acx.ACOUSTIC4)
    U = acx.field_gaussians(grid, 2025, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08))
    g = orc.grid3d(n_sample, grid.h, orc.ACOUSTIC4)
    dt = 0.4 * grid.h / 3.0
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps:
        orc.ts_solve("orc_rhs3d_cb_faithful", g, U, dt, dt, 0, 4)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": n_sample ** 3 * STAGES * steps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n_sample}^3 ACOUSTIC4 FP64, same 4-Gaussian workload, {steps} SSP-RK3 steps "
                      f"({STAGES * steps} RHS) of the oracle's reference-structured RHS, {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import acoustic_petsc_amd as acx

    N = args.n
    grid = acx.Grid(n=(N, N, N), h=1.0 / N, system=acx.ACOUSTIC4)
    U0 = acx.field_gaussians(grid, 2025, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08))
    U = torch.from_numpy(U0).cuda()
    work = (torch.empty_like(U), torch.empty_like(U))
    dt = acx.cfl_dt(grid, 0.4)
    if args.calib_copy:
        tmp = torch.empty_like(U)
        tmp.copy_(U)
        torch.cuda.synchronize()
        del tmp

    t_gpu, wall = time_steps(torch, lambda: acx.ssp_rks3_step(grid, U, dt, work), args.steps, args.warmup)
    cells = grid.ncells
    value = cells * STAGES * args.steps / t_gpu
    launch_s = t_gpu / (STAGES * args.steps)
    achieved = BYTES_FP64_AC4 * cells / launch_s / 1e9
    traffic = traffic_from_profiles(N)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_gpu / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D linearised acoustics CWENO4 (shipped linear weights) + LLF, {N}^3 periodic, "
                               f"FP64, SSP-RK3 4 stages, CFL 0.4, 4 seeded Gaussian pulses",
                   "grid": [N, N, N], "gamma": [1.0, 0.0], "precision": "fp64", "steps": args.steps,
                   "wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "k_linear<double, ACOUSTIC4> (fused RHS + stage)",
                     "bytes_per_cell_update": BYTES_FP64_AC4},
    }

    if args.variants == "all":
        var = {}
        # nonlinear CWENO4 FP64 on the same workload
        gnl = acx.Grid(n=(N, N, N), h=1.0 / N, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
        Unl = torch.from_numpy(U0).cuda()
        k = max(2, min(10, args.steps))
        t, _ = time_steps(torch, lambda: acx.ssp_rks3_step(gnl, Unl, dt, work), k, 1)
        v = cells * STAGES * k / t
        var["fp64_cweno_g085"] = {"value": v, "ms_per_step": 1e3 * t / k, "steps": k,
                                  "hbm_frac": v * BYTES_FP64_AC4 / 1e9 / HBM_PEAK_GBS}
        del Unl
        # FP32 compensated at 384^3 (band-limited noise + pulse), error vs FP64
        M = 384 if N == 256 else N
        g3 = acx.Grid(n=(M, M, M), h=1.0 / M, system=acx.ACOUSTIC4)
        V0 = acx.field_noise(g3, 3030, 4.0, 16.0, 1e-4, 1e-3, 0.05)
        dt3 = acx.cfl_dt(g3, 0.4)
        st = acx.CompensatedState(g3, torch.from_numpy(V0).cuda())
        k3 = max(2, min(20, args.steps))
        t, _ = time_steps(torch, lambda: st.step(dt3), k3, 1)
        v = M ** 3 * STAGES * k3 / t
        ref = torch.from_numpy(V0).cuda()
        w3 = (torch.empty_like(ref), torch.empty_like(ref))
        st_err = acx.CompensatedState(g3, ref)
        for _ in range(k3):
            acx.ssp_rks3_step(g3, ref, dt3, w3)
            st_err.step(dt3)
        rel, mx = acx.diff_f32c(st_err.hi, st_err.lo, ref)
        var["fp32c_linear_384"] = {"value": v, "ms_per_step": 1e3 * t / k3, "steps": k3,
                                   "bytes_per_cell_update": f32c_bytes_per_update(),
                                   "hbm_frac": v * f32c_bytes_per_update() / 1e9 / HBM_PEAK_GBS,
                                   "rel_l2_vs_fp64": rel, "max_abs_vs_fp64": mx,
                                   "speedup_vs_fp64_linear": v / value}
        del st, st_err, ref, w3
        out["variants"] = var

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(acx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
