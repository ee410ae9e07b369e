"""Point source and receiver on the WENO path (3d/weno4/point_source:
hype.c:14-183, rhs_function.c:337-346, source.c:11-17) through the
PETSc-compatible surface.

* On a small grid the fused and the generic TSSolve paths reproduce the
  FP64 oracle (state and every receiver sample of every RHS evaluation).
* At the reference's own configuration (100^3, h = 0.5, c = 250, dt = 5e-4,
  251 steps, source (50,50,50), receiver (57,57,57)) the receiver trace is
  checked against the closed-form Green's-function solution
  u(t) = Source(t - r/c) / (4 pi c^2 r) that results/green.py evaluates, and
  against the reference's CTCS receiver trace of the same problem
  (3d/ctcs/point_source/results/p.dat, tests/golden/ctcs_point_source_p.json).
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "tests", "drivers", "point_source_3d")


def run_driver(tmp_path, N, h, dt, nt, src, rcv, gamma, mode):
    rcv_file, state_file = tmp_path / f"rcv_{mode}.dat", tmp_path / f"state_{mode}.bin"
    args = [DRV, str(N), repr(h), repr(dt), str(nt), *map(str, src), *map(str, rcv), repr(gamma[0]), repr(gamma[1]),
            mode, str(rcv_file), str(state_file)]
    print("[case]", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    trace = np.loadtxt(rcv_file)
    raw = state_file.read_bytes()
    assert np.frombuffer(raw[:8], ">i4")[0] == 1211214
    state = np.frombuffer(raw[8:], ">f8").astype(np.float64).reshape(N, N, N, 2)
    return trace, state


@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_point_source_matches_oracle(tmp_path, orc, gamma):
    N, h, dt, nt, src, rcv = 24, 0.5, 5e-4, 40, (12, 11, 12), (15, 15, 14)
    U, tt, tu = orc.point_source_run(N, h, 250.0, dt, nt, src, rcv, gamma=gamma)
    assert np.abs(U).max() > 0
    for mode in ("fused", "generic"):
        trace, state = run_driver(tmp_path, N, h, dt, nt, src, rcv, gamma, mode)
        err = np.linalg.norm(state - U) / np.linalg.norm(U)
        print(f"[case] {mode} gamma={gamma}: state rel L2 {err:.2e}, {len(trace)} receiver samples", flush=True)
        assert err < 1e-12
        assert trace.shape == (len(tt), 2)
        assert np.max(np.abs(trace[:, 0] - tt)) < 1e-6 * dt          # printed with %e
        scale = np.max(np.abs(tu))
        assert scale > 0
        assert np.max(np.abs(trace[:, 1] - tu)) < 1e-11 * scale      # printed with %.12e


def test_green_function_and_ctcs_trace(tmp_path):
    N, h, dt, nt, c = 100, 0.5, 5e-4, 251, 250.0
    src, rcv = (50, 50, 50), (57, 57, 57)
    trace, _ = run_driver(tmp_path, N, h, dt, nt, src, rcv, (1.0, 0.0), "fused")
    # one sample per RHS evaluation: the direct call at t = 0, then 4 per step;
    # the first stage of step n samples the state at t = n dt
    assert len(trace) == 1 + 4 * nt
    t = trace[1::4, 0]
    u = trace[1::4, 1]
    assert np.allclose(t, dt * np.arange(nt), rtol=0, atol=1e-9)
    f0, t0 = 100.0, 4.0 / 100.0
    r = h * np.sqrt(3 * 7 ** 2)
    s = t - r / c
    exact = -2.0 * (s - t0) * f0 ** 2 * np.exp(-(f0 ** 2) * (s - t0) ** 2) / (4 * np.pi * c ** 2 * r)
    rel = np.linalg.norm(u - exact) / np.linalg.norm(exact)
    ip, ie = int(np.argmax(u)), int(np.argmax(exact))
    print(f"[case] Green: rel L2 {rel:.3f}, peak {u[ip]:.4e} at {t[ip]:.4f} s vs {exact[ie]:.4e} at {t[ie]:.4f} s",
          flush=True)
    assert rel < 0.02
    assert abs(u[ip] - exact[ie]) < 0.01 * exact[ie]
    assert ip == ie
    with open(os.path.join(ROOT, "tests", "golden", "ctcs_point_source_p.json")) as f:
        g = json.load(f)
    key_t = next(k for k in g if k.startswith("t"))
    key_p = next(k for k in g if k.startswith("p") and k != key_t)
    tc, pc = np.asarray(g[key_t], float), np.asarray(g[key_p], float)
    uc = np.interp(tc, t, u)
    relc = np.linalg.norm(uc - pc) / np.linalg.norm(pc)
    print(f"[case] vs CTCS p.dat: rel L2 {relc:.3f}", flush=True)
    assert relc < 0.02
