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

Beside the headline line's roofline: the read + write rate of state-sized
device copies on the same stream (the library's nontemporal VecCopy kernel
and torch's copy_), which bounds what a stage that streams the state can
reach on the box it runs on.

Also reported ("variants"): the nonlinear CWENO4 (gamma = 0.85/0.85) FP64
kernel on the same workload; the FP32 compensated path at 384^3 with its
error against FP64; the learned troubled-cell indicator's selective FP32
stepper at 256^3 and 384^3 (ReLU network, and the tanh network at 384^3)
against FP64 full CWENO4; the 48-problem ensemble; the 2D acoustic bubble
(eager and graph-replayed); the CTCS leapfrog sibling; the central WAVE2
system.  A CPU baseline (the C oracle on the host cores, a bounded sample)
is timed unless --no-cpu.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
                       [--variants all|none|name,...] [--learned-sizes 256,384]
                       [--no-cpu] [--calib-copy]
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
    ap.add_argument("--variants", default="all",
                    help="all, none, or a comma list of fp64_cweno, fp32c, learned, ensemble")
    ap.add_argument("--learned-sizes", default="256,384",
                    type=lambda s: [int(v) for v in s.split(",") if v])
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


def time_steps(torch, step, steps, warmup, per_step=None):
    """GPU time of `steps` calls of step() after `warmup` untimed ones, from
    HIP events on the launch stream (torch's current stream, on which the
    library launches).  With a dict `per_step`, an event is also recorded
    between consecutive steps and the median / min / max step time (s) are
    stored there."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1 if per_step is not None else 2)]
    t0 = time.perf_counter()
    ev[0].record(s)
    for i in range(steps):
        step()
        if per_step is not None:
            ev[i + 1].record(s)
    ev[-1].record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if per_step is not None and steps > 0:
        dts = sorted(ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(steps))
        per_step.update({"median_s": dts[len(dts) // 2], "min_s": dts[0], "max_s": dts[-1]})
    return ev[0].elapsed_time(ev[-1]) / 1e3, wall


def settle(torch, step, batch=10, tol=0.01, max_s=3.0):
    """Runs untimed batches of `batch` steps until the GPU time of two
    consecutive batches agrees within `tol` (clocks and caches at steady
    state), at most `max_s` seconds.  Returns (steps run, last batch times)."""
    n, last, t_end = 0, None, time.perf_counter() + max_s
    while True:
        t, _ = time_steps(torch, step, batch, 0)
        n += batch
        if last is not None and abs(t - last) <= tol * min(t, last):
            return n, (last, t)
        if time.perf_counter() > t_end:
            return n, (last, t)
        last = t


def note(msg):
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def copy_ceiling(torch, acx, U, W, reps=20):
    """Read + write rate (GB/s) of state-sized device copies, timed like the
    stages on the launch stream: the library's nontemporal VecCopy kernel
    (acx_vec_copy_f64) and torch's copy_.  A stage kernel that reads and
    writes the state once cannot beat the faster of the two by much."""
    nbytes = 2 * U.numel() * U.element_size()
    t_acx, _ = time_steps(torch, lambda: acx.vec_copy(U, W), reps, 5)
    t_torch, _ = time_steps(torch, lambda: W.copy_(U), reps, 5)
    return {"bytes_per_copy": nbytes, "acx_vec_copy_GBs": nbytes * reps / t_acx / 1e9,
            "torch_copy_GBs": nbytes * reps / t_torch / 1e9}


def _host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(acx, n_sample=96, max_steps=24, budget_s=10.0):
    """The FP64 oracle (reference loop structure: ghosted copy, per-Gauss-point
    scratch, per-Gauss-point LLF) timed on this host on a 96^3 sample of the
    same workload, on all available cores (up to 16) and on one core."""
    import numpy as np
    from oracle import oracle as orc
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    grid = acx.Grid(n=(n_sample,) * 3, h=1.0 / n_sample, system=acx.ACOUSTIC4)
    U0 = acx.field_gaussians(grid, 2025, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08))
    g = orc.grid3d(n_sample, grid.h, orc.ACOUSTIC4)
    dt = 0.4 * grid.h / 3.0

    def run(threads, steps_cap, budget, Uinit, gg, n, cb="orc_rhs3d_cb_faithful"):
        orc.lib().orc_set_threads(threads)
        U = Uinit.copy()
        steps, t0 = 0, time.perf_counter()
        while steps < steps_cap:
            orc.ts_solve(cb, gg, U, dt * n_sample / n, dt * n_sample / n, 0, 4)
            steps += 1
            if time.perf_counter() - t0 >= budget:
                break
        el = time.perf_counter() - t0
        return n ** 3 * STAGES * steps / el, steps, el

    threads = max(1, min(avail, 16))
    v, steps, el = run(threads, max_steps, budget_s, U0, g, n_sample)
    # one core on a 64^3 sample of the same workload (a 96^3 step takes ~10 s on one core)
    n1 = 64
    grid1 = acx.Grid(n=(n1,) * 3, h=1.0 / n1, system=acx.ACOUSTIC4)
    U1 = acx.field_gaussians(grid1, 2025, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08))
    v1, steps1, el1 = run(1, 3, budget_s, U1, orc.grid3d(n1, grid1.h, orc.ACOUSTIC4), n1)
    # the oracle's streaming face-average form (no scratch, one pass per
    # stage, OpenMP over z-slabs): what a tuned CPU code of the same scheme does
    vs, steps_s, el_s = run(threads, 4 * max_steps, budget_s / 2, U0, g, n_sample, "orc_rhs3d_cb_stream")
    # the 2D case of the reference's CPU configuration (256^2 linearised
    # acoustics, CFL 0.4): the oracle's 2D restatement, one core, bounded
    N2 = 256
    U2 = np.ascontiguousarray(pulse_2d(N2)[0, :, :, :3])
    g2 = orc.G2((N2, N2), 1.0 / N2, orc.ACOUSTIC4, 0, 1.0, 0.0, 1.0, 1.0, (0, 0))
    dt2 = 0.4 * (1.0 / N2) / 2.0
    steps2, t2 = 0, time.perf_counter()
    while steps2 < 200 and time.perf_counter() - t2 < budget_s / 2:
        orc.ts_solve("orc_rhs2d_cb", g2, U2, dt2, dt2, 0, 4)
        steps2 += 1
    el2 = time.perf_counter() - t2
    return {"value": v, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n_sample}^3 ACOUSTIC4 FP64, same 4-Gaussian workload, {steps} SSP-RK3 steps "
                      f"({STAGES * steps} RHS) of the oracle's reference-structured RHS, {el:.1f} s",
            "single_core": {"value": v1, "sample": f"{n1}^3, {steps1} steps, {el1:.1f} s"},
            "streaming": {"value": vs, "cores": threads,
                          "sample": f"{n_sample}^3, {steps_s} steps of the streaming face-average RHS, {el_s:.1f} s"},
            "acoustic2d_256": {"value": N2 * N2 * STAGES * steps2 / el2, "cores": 1,
                               "sample": f"256^2 2D linearised acoustics, {steps2} of 200 steps, {el2:.1f} s"},
            "host": {"cpu": _host_cpu_model(), "available_cores": avail, "nproc": os.cpu_count()}}


VARIANTS = ("fp64_cweno", "fp32c", "learned", "ensemble", "2d", "ctcs", "wave2")


def pulse_2d(N, seed=1001, amp=1e-3, sigma=0.05):
    """The 2D linearised-acoustics initial state on the periodic unit square,
    N^2 cells: p' = amp exp(-r^2 / (2 sigma^2)) about a seeded centre (minimum
    image), as exact cell averages (a product of erf differences, nine
    periodic images), u = v = 0.  Returned as (1, N, N, 4) with w = 0: the
    one-plane ACOUSTIC4 grid on which the library runs 2D programs."""
    import numpy as np
    from scipy.special import erf
    c = np.random.default_rng(seed).random(2)
    h = 1.0 / N
    edges = np.arange(N + 1) * h
    k = sigma * np.sqrt(np.pi / 2.0) / h
    prof = []
    for d in range(2):
        acc = np.zeros(N)
        for img in (-1.0, 0.0, 1.0):
            z = (edges - c[d] + img) / (sigma * np.sqrt(2.0))
            acc += erf(z[1:]) - erf(z[:-1])
        prof.append(k * acc)
    U = np.zeros((1, N, N, 4))
    U[0, :, :, 0] = amp * np.outer(prof[1], prof[0])
    return U


def variant_2d(torch, acx, steps=200, N=256):
    """The 2D case of the reference's CPU configuration: linearised acoustics
    (p', u, v; rho0 = c0 = 1) on a 256^2 periodic grid, FP64, SSP-RK3 at CFL
    0.4 (dt = CFL h / 2), 200 steps, one Gaussian pulse, run on the GPU as a
    one-plane grid."""
    g = acx.Grid(n=(N, N, 1), h=1.0 / N, system=acx.ACOUSTIC4)
    U = torch.from_numpy(pulse_2d(N)).cuda()
    work = (torch.empty_like(U), torch.empty_like(U))
    dt = acx.cfl_dt(g, 0.4)
    t, _ = time_steps(torch, lambda: acx.ssp_rks3_step(g, U, dt, work), steps, 2)
    # the same steps replayed from the captured hipGraph of one step
    per = 20
    tg, _ = time_steps(torch, lambda: acx.ssp_rks3_steps_graph(g, U, dt, per, work), steps // per, 1)
    return {"value": N * N * STAGES * steps / tg, "ms_per_step": 1e3 * tg / steps, "steps": steps,
            "grid": [N, N, 1], "stepper": f"hipGraph replays of {per} captured steps, one launch per step",
            "eager_value": N * N * STAGES * steps / t, "eager_ms_per_step": 1e3 * t / steps,
            "note": "one launch per step (16 x 16 tiles with a 12-cell halo, the four stages in LDS); "
                    "eager steps add a copy back into U"}


def variant_wave2(torch, acx, N=256, steps=20):
    """The literal 3d/weno4/plane_wave system (u, v; du/dt = v, dv/dt = c^2
    div grad u) at N^3, FP64, SSP-RK3 at CFL 0.4, 4 seeded Gaussian pulses in
    u: the shipped linear weights (HBM-bound, 9/4 x 2 x 8 = 36 B per
    cell-update) and the nonlinear 0.85/0.85 weights."""
    out = {}
    for name, gam in (("linear", (1.0, 0.0)), ("nonlinear_g085", (0.85, 0.85))):
        g = acx.Grid(n=(N, N, N), h=1.0 / N, system=acx.WAVE2, gamma_hi=gam[0], gamma_lo=gam[1])
        U = torch.from_numpy(acx.field_gaussians(g, 2025)).cuda()
        work = (torch.empty_like(U), torch.empty_like(U))
        dt = acx.cfl_dt(g, 0.4)
        k = steps if name == "linear" else max(2, steps // 4)
        t, _ = time_steps(torch, lambda: acx.ssp_rks3_step(g, U, dt, work), k, 1)
        v = g.ncells * STAGES * k / t
        out[name] = {"value": v, "ms_per_step": 1e3 * t / k, "steps": k, "grid": [N, N, N],
                     "bytes_per_cell_update": 36, "hbm_frac": v * 36 / 1e9 / HBM_PEAK_GBS}
        del U, work
        torch.cuda.empty_cache()
    return out


def variant_ctcs(torch, acx, steps=100):
    """The CTCS leapfrog sibling (3d/ctcs/point_source/wave.c:138-215): 4th-order
    13-point Laplacian + point source, FP64, on the reference's 100^3 problem
    and at 256^3 (same parameters, source and receiver at the centre).  One
    update reads Pold, P and the speed C and writes Pnew: 32 B per cell."""
    out = {}
    for n in (100, 256):
        p = acx.Ctcs()
        p.n, p.stencil = (n, n, n), 2   # stencil width 2: the 4th-order 13-point Laplacian
        p.src = p.rcv = (n // 2, n // 2, n // 2)
        cs = p.c_struct()
        shape = (n, n, n)
        Pold = torch.zeros(shape, dtype=torch.float64, device="cuda")
        P, Pnew = torch.zeros_like(Pold), torch.zeros_like(Pold)
        C = torch.full_like(Pold, p.c0)
        bufs = [Pold, P, Pnew]
        it = [0]

        def step():
            a, b, c = bufs
            acx._check(acx.lib().acx_ctcs_update_f64(acx._C.byref(cs), acx._ptr(a, torch.float64),
                                                     acx._ptr(b, torch.float64), acx._ptr(c, torch.float64),
                                                     acx._ptr(C, torch.float64), it[0], None, acx._stream()))
            bufs[:] = [b, c, a]
            it[0] += 1

        t, _ = time_steps(torch, step, steps, 5)
        v = n ** 3 * steps / t
        out[f"n{n}"] = {"value": v, "ms_per_step": 1e3 * t / steps, "steps": steps, "grid": [n, n, n],
                        "bytes_per_cell_update": 32, "hbm_frac": v * 32 / 1e9 / HBM_PEAK_GBS}
        del Pold, P, Pnew, C, bufs
        torch.cuda.empty_cache()
    return out


def variant_fp64_cweno(torch, acx, N, U0, work, dt, steps):
    """Nonlinear CWENO4 (gamma = 0.85/0.85) FP64 on the headline workload."""
    g = acx.Grid(n=(N, N, N), h=1.0 / N, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    U = torch.from_numpy(U0).cuda()
    k = max(2, min(10, steps))
    t, _ = time_steps(torch, lambda: acx.ssp_rks3_step(g, U, dt, work), k, 1)
    v = g.ncells * STAGES * k / t
    return {"value": v, "ms_per_step": 1e3 * t / k, "steps": k,
            "hbm_frac": v * BYTES_FP64_AC4 / 1e9 / HBM_PEAK_GBS}


def variant_fp32c(torch, acx, M, steps, fp64_value, err_steps=200):
    """FP32 storage + compensated accumulation, shipped linear weights, M^3,
    band-limited noise (|k| in [4, 16], RMS 1e-4) + one Gaussian pulse; error
    against the FP64 stepper (equal to the FP64 oracle to 1e-12) after
    `err_steps` steps."""
    g = acx.Grid(n=(M, M, M), h=1.0 / M, system=acx.ACOUSTIC4)
    V0 = acx.field_noise(g, 3030, 4.0, 16.0, 1e-4, 1e-3, 0.05)
    dt = acx.cfl_dt(g, 0.4)
    st = acx.CompensatedState(g, torch.from_numpy(V0).cuda())
    k = max(2, min(20, steps))
    t, _ = time_steps(torch, lambda: st.step(dt), k, 1)
    v = M ** 3 * STAGES * k / t
    del st
    ref = torch.from_numpy(V0).cuda()
    w = (torch.empty_like(ref), torch.empty_like(ref))
    st = acx.CompensatedState(g, ref)
    for _ in range(err_steps):
        acx.ssp_rks3_step(g, ref, dt, w)
        st.step(dt)
    rel, mx = acx.diff_f32c(st.hi, st.lo, ref)
    out = {"value": v, "ms_per_step": 1e3 * t / k, "steps": k, "grid": [M, M, M],
           "bytes_per_cell_update": f32c_bytes_per_update(),
           "hbm_frac": v * f32c_bytes_per_update() / 1e9 / HBM_PEAK_GBS,
           "error_steps": err_steps, "rel_l2_vs_fp64": rel, "max_abs_vs_fp64": mx,
           "max_abs_state": float(ref.abs().max()),
           "speedup_vs_fp64_linear_256": v / fp64_value}
    del st, ref, w
    torch.cuda.empty_cache()
    return out


def train_indicator(torch, acx, acts=("relu",)):
    """The troubled-cell MLP trained on the GPU by the library's own kernels:
    8 solver-labelled snapshots of 64^3 steep-pulse runs (2.1e6 cells, 8.4e6
    (cell, component) samples), labelled positive where the eight normalised
    CWENO4 weights leave the linear weights, sum_k |w_k - gamma_k| > 0.05.
    One network per hidden activation in `acts`, on the same samples."""
    from acoustic_petsc_amd import learn
    t0 = time.perf_counter()
    X, T = learn.make_dataset(n_snapshots=8, N=64, seed0=100)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    out = {}
    for act in acts:
        t0 = time.perf_counter()
        p = learn.train(X, T, epochs=6, batch=8192, lr=3e-3, seed=1, act=act)
        packed = learn.pack(p, act=act)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rec, prec, frac = learn.evaluate(packed, X, T, 0.3)
        out[act] = (packed, {"activation": act, "samples": int(X.shape[0]),
                             "positive_frac": float(T.float().mean()), "dataset_s": t_data, "train_s": el,
                             "recall": rec, "precision": prec, "flagged_frac": frac})
    del X, T
    return out


def variant_learned(torch, acx, sizes, steps, var, fp64_value):
    """FP32 compensated stepper with the learned indicator (full CWENO4 on
    flagged cells, linear reconstruction elsewhere) against the FP64 nonlinear
    stepper on a held-out steep pulse, nonlinear weights gamma = 0.85/0.85."""
    from acoustic_petsc_amd import learn
    nets = train_indicator(torch, acx, ("relu", "tanh"))
    packed, info = nets["relu"]
    res = {"indicator_training": info}
    for M in sizes:
        g = acx.Grid(n=(M, M, M), h=1.0 / M, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
        glin = acx.Grid(n=(M, M, M), h=1.0 / M, system=acx.ACOUSTIC4)
        U0 = torch.from_numpy(learn.steep_pulse_field(g, 4242))
        dt = acx.cfl_dt(g, 0.4)
        k = max(2, min(20, steps))
        cells = M ** 3
        # FP64 all-WENO: the error reference and the plain FP64 throughput at M^3
        ref = U0.cuda()
        w = (torch.empty_like(ref), torch.empty_like(ref))
        acx.ssp_rks3_step(g, ref, dt, w)     # warm-up step, mirrored below
        t64, _ = time_steps(torch, lambda: acx.ssp_rks3_step(g, ref, dt, w), k, 0)
        lin64 = U0.cuda()
        tl64, _ = time_steps(torch, lambda: acx.ssp_rks3_step(glin, lin64, dt, w), k, 1)
        del lin64, w
        sel = learn.SelectiveStepper(g, packed, U0.cuda(), thr=0.3, refresh=12, dilate=3)
        sel.step(dt)
        fr = []
        tsel, _ = time_steps(torch, lambda: sel.step(dt), k, 0)
        fr.append(sel.flagged_fraction())
        rel, mx = acx.diff_f32c(sel.state.hi, sel.state.lo, ref)
        # cost of one full-WENO judge pass and of one indicator pass on this grid
        t_judge, _ = time_steps(torch, lambda: learn.judge(g, sel.state.hi, used=sel.flags, base=sel.tmp,
                                                           out=sel.tmp2, tol=sel.judge_tol), 3, 1)
        t_ind, _ = time_steps(torch, lambda: sel._mlp_flags(), 3, 1)
        judge_info = {"windows_judged": sel.judged_windows, "misses": sel.misses,
                      "windows_rerun": sel.redone_windows, "judge_ms": 1e3 * t_judge / 3,
                      "indicator_ms": 1e3 * t_ind / 3, "tolerance": sel.judge_tol,
                      "largest_skipped_face_difference": sel.worst_skipped,
                      "rule": "every 12-step window: at each (cell, component) run with the linear polynomial, "
                              "full CWENO4 vs linear face averages of the state reached; a difference above "
                              "1e-6 max|U| reruns the window with full CWENO4"}
        del sel
        lin = acx.CompensatedState(glin, U0.cuda())
        for _ in range(k + 1):
            lin.step(dt)
        rel_lin, _ = acx.diff_f32c(lin.hi, lin.lo, ref)
        del lin
        full = acx.CompensatedState(g, U0.cuda())
        full.step(dt)
        tfull, _ = time_steps(torch, lambda: full.step(dt), max(2, k // 4), 0)
        del full, ref
        torch.cuda.empty_cache()
        v = cells * STAGES * k / tsel
        v64 = cells * STAGES * k / t64
        vl64 = cells * STAGES * k / tl64
        res[f"fp32_learned_{M}"] = {
            "value": v, "ms_per_step": 1e3 * tsel / k, "steps": k + 1, "grid": [M, M, M],
            "indicator": {"threshold": 0.3, "refresh_steps": 12, "dilation_cells": 3, "label": "sum|w-gamma| > 0.05"},
            "judge": judge_info,
            "flagged_frac_last_step": fr[-1], "skipped_frac_last_step": 1.0 - fr[-1],
            "rel_l2_vs_fp64_weno": rel, "max_abs_vs_fp64_weno": mx,
            "linear_only_rel_l2_vs_fp64_weno": rel_lin,
            "fp64_weno_value": v64, "fp64_linear_value": vl64,
            "fp32_all_weno_value": cells * STAGES * max(2, k // 4) / tfull,
            "speedup_vs_fp64_weno": v / v64, "speedup_vs_fp64_linear": v / vl64}
    # the same stepper with the tanh network, largest grid only
    M = sizes[-1]
    g = acx.Grid(n=(M, M, M), h=1.0 / M, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    U0 = torch.from_numpy(learn.steep_pulse_field(g, 4242))
    dt = acx.cfl_dt(g, 0.4)
    k = max(2, min(20, steps))
    ref = U0.cuda()
    w = (torch.empty_like(ref), torch.empty_like(ref))
    for _ in range(k + 1):
        acx.ssp_rks3_step(g, ref, dt, w)
    del w
    tpacked, tinfo = nets["tanh"]
    sel = learn.SelectiveStepper(g, tpacked, U0.cuda(), thr=0.3, refresh=12, dilate=3)
    sel.step(dt)
    tsel, _ = time_steps(torch, lambda: sel.step(dt), k, 0)
    rel, mx = acx.diff_f32c(sel.state.hi, sel.state.lo, ref)
    t_ind, _ = time_steps(torch, lambda: sel._mlp_flags(), 3, 1)
    v = M ** 3 * STAGES * k / tsel
    res["tanh_activation"] = {
        "training": tinfo, "grid": [M, M, M], "value": v, "ms_per_step": 1e3 * tsel / k,
        "indicator_ms": 1e3 * t_ind / 3, "flagged_frac_last_step": sel.flagged_fraction(),
        "rel_l2_vs_fp64_weno": rel, "max_abs_vs_fp64_weno": mx,
        "judge": {"windows_judged": sel.judged_windows, "misses": sel.misses, "windows_rerun": sel.redone_windows},
        "relu_value": res[f"fp32_learned_{M}"]["value"]}
    del sel, ref
    torch.cuda.empty_cache()
    return res


def variant_ensemble(torch, acx, steps, nprob=48, seed=4848):
    """48 independent ACOUSTIC4 problems, edges from {64, 96, 128, 160} per
    axis, 1-3 Gaussian pulses each, FP32 compensated, persistent work-queue
    kernel + captured two-step hipGraph, autotuned tile configuration; against
    48 sequential single-problem FP32 steppers, and per-problem error against
    the FP64 stepper."""
    from acoustic_petsc_amd.ensemble import Ensemble, random_dims
    dims = random_dims(nprob, seed)
    h = 1.0 / 128
    g0 = acx.Grid(n=(64, 64, 64), h=h, system=acx.ACOUSTIC4)
    dt = acx.cfl_dt(g0, 0.4)
    fields = [acx.field_gaussians(acx.Grid(n=d, h=h, system=acx.ACOUSTIC4), seed + p, npulse=1 + p % 3,
                                  amp=(1e-4, 1e-3), sigma=(0.05, 0.15)) for p, d in enumerate(dims)]
    e = Ensemble(dims, h)
    e.load(fields)
    best, ms_best = e.autotune(dt)
    k = max(2, min(100, steps))
    k += k % 2
    t_ens, _ = time_steps(torch, lambda: e.graph_steps(dt, 2), k // 2, 1)
    cells = e.ncells
    v_ens = cells * STAGES * k / t_ens
    # sequential single-problem launches
    states = [acx.CompensatedState(acx.Grid(n=d, h=h, system=acx.ACOUSTIC4), torch.from_numpy(f).cuda())
              for d, f in zip(dims, fields)]

    def seq_step():
        for s in states:
            s.step(dt)
    t_seq, _ = time_steps(torch, seq_step, k, 1)
    v_seq = cells * STAGES * k / t_seq
    # correctness: 10 steps from the initial fields, ensemble vs single-problem vs FP64
    ne = 10
    e.load(fields)
    e.graph_steps(dt, ne)
    hi, lo = e.state32()
    hi, lo = hi.cpu().numpy(), lo.cpu().numpy()
    bitwise, worst = True, 0.0
    for p, (d, f) in enumerate(zip(dims, fields)):
        g = acx.Grid(n=d, h=h, system=acx.ACOUSTIC4)
        s = acx.CompensatedState(g, torch.from_numpy(f).cuda())
        ref = torch.from_numpy(f).cuda()
        w = (torch.empty_like(ref), torch.empty_like(ref))
        for _ in range(ne):
            s.step(dt)
            acx.ssp_rks3_step(g, ref, dt, w)
        a, b = 4 * e.offsets[p], 4 * e.offsets[p + 1]
        bitwise &= bool((s.hi.cpu().numpy().ravel() == hi[a:b]).all() and (s.lo.cpu().numpy().ravel() == lo[a:b]).all())
        rel, _ = acx.diff_f32c(torch.from_numpy(hi[a:b]).cuda(), torch.from_numpy(lo[a:b]).cuda(), ref.reshape(-1))
        worst = max(worst, rel)
    del states
    e.close()
    torch.cuda.empty_cache()
    return {"value": v_ens, "ms_per_step": 1e3 * t_ens / k, "steps": k, "problems": nprob,
            "cells": cells, "autotuned_config": best, "autotune_ms_per_step": ms_best,
            "sequential_value": v_seq, "speedup_vs_sequential": v_ens / v_seq,
            "hbm_frac": v_ens * f32c_bytes_per_update() / 1e9 / HBM_PEAK_GBS,
            "check_steps": ne, "bitwise_equal_single_problem": bitwise,
            "max_rel_l2_vs_fp64": worst}


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

    headline_step = lambda: acx.ssp_rks3_step(grid, U, dt, work)
    note(f"headline {N}^3: settling")
    settle_steps, settle_pair = settle(torch, headline_step)
    per_step = {}
    t_gpu, wall = time_steps(torch, headline_step, args.steps, args.warmup, per_step)
    cells = grid.ncells
    value = cells * STAGES * args.steps / t_gpu
    launch_s = t_gpu / (STAGES * args.steps)
    achieved = BYTES_FP64_AC4 * cells / launch_s / 1e9
    note(f"headline {N}^3: {1e3 * t_gpu / args.steps:.4f} ms per step after {settle_steps} settle steps")
    traffic = traffic_from_profiles(N)
    copy = copy_ceiling(torch, acx, U, work[0])
    note(f"copy ceiling: {copy['acx_vec_copy_GBs']:.0f} GB/s (acx_vec_copy), {copy['torch_copy_GBs']:.0f} GB/s (torch)")
    copy_peak = max(copy["acx_vec_copy_GBs"], copy["torch_copy_GBs"])
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
        "timing": {"settle_steps": settle_steps,
                   "settle_last_batches_ms": [1e3 * x if x is not None else None for x in settle_pair],
                   "step_ms_median": 1e3 * per_step.get("median_s", float("nan")),
                   "step_ms_min": 1e3 * per_step.get("min_s", float("nan")),
                   "step_ms_max": 1e3 * per_step.get("max_s", float("nan")),
                   "note": "untimed batches of 10 steps run until two consecutive batches agree "
                           "within 1 %, then the warmup steps, then the timed steps"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": f"profiles/k_linear_fp64_ac4_{N}_pmc.md (FETCH_SIZE/WRITE_SIZE "
                                       "of a rocprofv3 --pmc run, not measured in this run)"
                     if traffic is not None else None,
                     "copy_ceiling": copy,
                     "frac_of_copy": achieved / copy_peak,
                     "kernel": "k_linear<double, ACOUSTIC4> (fused RHS + stage)",
                     "bytes_per_cell_update": BYTES_FP64_AC4,
                     "bytes_per_launch": BYTES_FP64_AC4 * cells},
    }

    if args.variants != "none":
        want = set(VARIANTS) if args.variants == "all" else set(args.variants.split(","))
        unknown = want - set(VARIANTS)
        if unknown:
            raise SystemExit(f"unknown variants {sorted(unknown)}; choose from {VARIANTS}")
        var = {}
        if "fp64_cweno" in want:
            note("variant fp64_cweno")
            var["fp64_cweno_g085"] = variant_fp64_cweno(torch, acx, N, U0, work, dt, args.steps)
        if "fp32c" in want:
            note("variant fp32c")
            var["fp32c_linear_384"] = variant_fp32c(torch, acx, 384 if N == 256 else N, args.steps, value)
        del work, U
        torch.cuda.empty_cache()
        if "learned" in want:
            note("variant learned")
            var.update(variant_learned(torch, acx, args.learned_sizes, args.steps, var, value))
        if "ensemble" in want:
            note("variant ensemble")
            var["ensemble_48"] = variant_ensemble(torch, acx, args.steps)
        if "2d" in want:
            note("variant 2d")
            var["acoustic2d_256"] = variant_2d(torch, acx)
        if "ctcs" in want:
            note("variant ctcs")
            var["ctcs_leapfrog"] = variant_ctcs(torch, acx)
        if "wave2" in want:
            note("variant wave2")
            var["wave2_256"] = variant_wave2(torch, acx)
        out["variants"] = var

    if not args.no_cpu:
        note("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(acx)
    note("done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
