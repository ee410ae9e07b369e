"""2D siblings (2d/weno4/standing_wave, config-1 style 2D linearised acoustics)
on the GPU as one-plane grids, n = (Nx, Ny, 1) with a periodic z axis.

* RHS and SSP-RK3 steps against the 2D oracle (`orc_rhs2d_faithful`, the
  restatement of 2d/weno4/standing_wave/rhs_function.c:12-257 and of the 2D
  LLF acoustics) at 1e-12: WAVE2 periodic and adiabatic walls; ACOUSTIC4
  with w = 0 against the 3-variable (p', u, v) 2D operator, w staying 0.
* The FP32 compensated stepper on a 2D grid against the FP64 stepper.
* The reference 2D program (2d/weno4/standing_wave/hype.c, TSSSP RK104,
  adiabatic walls, T = 1) through petsclite reproduces
  results/Error.dat (tests/golden/standing_wave_2d_error.json).
* Nonlinear weights: the 2D CWENO4 of 2d/weno4/bubble/reconstruction.c:128-251
  (five quadratic stencils) against the oracle's 2D CWENO4 (orc_cweno2d,
  pinned by the stencil-footprint and jump-selection checks), RHS and steps
  at 1e-12, FP32 compensated against FP64.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def oracle_rhs2d(orc, N, h, system, bc, U2, gamma=(1.0, 0.0)):
    nonlinear = 0 if gamma == (1.0, 0.0) else 1
    g2 = orc.G2((N[0], N[1]), h, system, nonlinear, gamma[0], gamma[1], 1.0, 1.0, bc)
    R = np.zeros_like(U2)
    assert orc.lib().orc_rhs2d_faithful(C.byref(g2), orc.P(U2), orc.P(R)) == 0
    return R, g2


@pytest.mark.parametrize("N", [(12, 9), (16, 16), (70, 33)])
def test_one_launch_2d_step_matches_the_oracle(N, acx, orc, torch_gpu):
    """Periodic 2D linearised acoustics with the shipped weights steps in one
    launch (tiles of 16 x 16 with a 12-cell halo, the four stages in LDS):
    against the oracle's 2D RKS3 steps at 1e-12 on grids smaller than the
    halo (wrapping several times), exactly one tile, and partial tiles; the
    graph-replayed steps (alternating between U and W0, an odd count ending
    with a copy) equal the eager ones bit for bit."""
    torch = torch_gpu
    h = 1.0 / N[0]
    rng = np.random.default_rng(N[0] * 100 + N[1])
    U2 = rng.standard_normal((N[1], N[0], 3))
    U = np.zeros((1, N[1], N[0], 4))
    U[0, ..., :3] = U2
    grid = acx.Grid(n=(N[0], N[1], 1), h=h, system=1)
    dt = acx.cfl_dt(grid, 0.4)
    g2 = orc.G2((N[0], N[1]), h, 1, 0, 1.0, 0.0, 1.0, 1.0, (0, 0))
    ref = U2.copy()
    for _ in range(3):
        orc.ts_solve("orc_rhs2d_cb", g2, ref, dt, dt, 0, 4)
    print(f"[case] one-launch 2d step n={N}", flush=True)
    a = torch.from_numpy(U).cuda()
    b = a.clone()
    wa = (torch.empty_like(a), torch.empty_like(a))
    wb = (torch.empty_like(b), torch.empty_like(b))
    for _ in range(3):
        acx.ssp_rks3_step(grid, a, dt, wa)
    acx.ssp_rks3_steps_graph(grid, b, dt, 3, wb)
    got = a.cpu().numpy()
    assert rel_l2(got[0, ..., :3], ref) <= TOL
    assert np.all(got[..., 3] == 0.0)
    assert torch.equal(a, b)


@pytest.mark.parametrize("system,bc", [(0, (0, 0)), (0, (1, 1)), (0, (1, 0)), (1, (0, 0))],
                         ids=["wave2-periodic", "wave2-adiabatic", "wave2-mixed", "acoustic-periodic"])
def test_rhs_and_steps_match_2d_oracle(system, bc, acx, orc, torch_gpu):
    torch = torch_gpu
    N, h = (40, 24), 1.0 / 40
    nv2 = 2 if system == 0 else 3
    rng = np.random.default_rng(11 + system)
    U2 = rng.standard_normal((N[1], N[0], nv2))
    U = np.zeros((1, N[1], N[0], 2 if system == 0 else 4))
    U[0, ..., :nv2] = U2
    grid = acx.Grid(n=(N[0], N[1], 1), h=h, system=system, bc=(bc[0], bc[1], 0))
    print(f"[case] 2d rhs system={system} bc={bc} n={N}", flush=True)
    R2, g2 = oracle_rhs2d(orc, N, h, system, bc, U2)
    got = acx.rhs(grid, torch.from_numpy(U).cuda()).cpu().numpy()
    assert rel_l2(got[0, ..., :nv2], R2) <= TOL
    if system == 1:
        assert np.all(got[..., 3] == 0.0)
    # three SSP-RK3 (s = 4) steps
    dt = acx.cfl_dt(grid, 0.4)
    assert dt == pytest.approx(0.4 * h / 2.0, rel=1e-15)      # D = 2 on a 2D grid
    ref = U2.copy()
    for _ in range(3):
        orc.ts_solve("orc_rhs2d_cb", g2, ref, dt, dt, 0, 4)
    Ud = torch.from_numpy(U).cuda()
    for _ in range(3):
        acx.ssp_rks3_step(grid, Ud, dt)
    got = Ud.cpu().numpy()
    err = rel_l2(got[0, ..., :nv2], ref)
    print(f"[case] 2d steps system={system} bc={bc} rel_l2={err:.3e}", flush=True)
    assert err <= TOL
    if system == 1:
        assert np.all(got[..., 3] == 0.0)


def test_fp32_compensated_on_2d_grid(acx, torch_gpu):
    torch = torch_gpu
    N = 96
    grid = acx.Grid(n=(N, N, 1), h=1.0 / N, system=acx.ACOUSTIC4)
    U = acx.field_gaussians(grid, 2025, npulse=3, amp=(1e-4, 1e-3), sigma=(0.05, 0.1))
    U[..., 3] = 0.0
    dt = acx.cfl_dt(grid, 0.4)
    print(f"[case] 2d f32c n={N}", flush=True)
    ref = torch.from_numpy(U).cuda()
    st = acx.CompensatedState(grid, torch.from_numpy(U).cuda())
    for _ in range(50):
        acx.ssp_rks3_step(grid, ref, dt)
        st.step(dt)
    rel, mx = acx.diff_f32c(st.hi, st.lo, ref)
    print(f"[case] 2d f32c rel_l2={rel:.2e} max={mx:.2e}", flush=True)
    assert rel < 1e-6


def _field2d(seed, N, nv2, kind):
    rng = np.random.default_rng(seed)
    if kind == "rough":
        return rng.standard_normal((N[1], N[0], nv2))
    # a disc with a jump plus noise: the weights leave the linear weights along the rim
    y, x = np.meshgrid(np.arange(N[1]), np.arange(N[0]), indexing="ij")
    inside = (x - N[0] / 2) ** 2 + 1.4 * (y - N[1] / 2) ** 2 < (min(N) / 3) ** 2
    U = np.zeros((N[1], N[0], nv2))
    for v in range(nv2):
        U[..., v] = np.where(inside, 1.0 + 0.3 * v, 0.2) + 0.01 * rng.standard_normal((N[1], N[0]))
    return U


@pytest.mark.parametrize("system,bc,gamma,kind", [
    (0, (0, 0), (0.85, 0.85), "rough"), (0, (1, 1), (0.85, 0.85), "disc"), (0, (1, 0), (0.85, 0.85), "rough"),
    (1, (0, 0), (0.85, 0.85), "disc"), (1, (0, 0), (0.85, 0.85), "rough"), (1, (0, 0), (0.5, 0.3), "rough"),
    (0, (0, 0), (0.9, 0.1), "disc")],
    ids=["wave2-per", "wave2-adia-disc", "wave2-mixed", "acoustic-disc", "acoustic-rough", "acoustic-g05-03",
         "wave2-g09-01"])
def test_2d_cweno4_matches_2d_oracle(system, bc, gamma, kind, acx, orc, torch_gpu):
    """The 2D CWENO4 (2d/weno4/bubble/reconstruction.c:128-251) on one-plane
    grids: RHS and three SSP-RK3 steps against the oracle's 2D RHS with
    nonlinear weights.  gamma = (0.5, 0.3) makes the central linear weight
    of the 3D form negative (its share of the two z stencils exceeds it):
    only the sum enters, and the result still matches."""
    torch = torch_gpu
    N, h = (36, 28), 1.0 / 36
    nv2 = 2 if system == 0 else 3
    U2 = _field2d(21 + system, N, nv2, kind)
    U = np.zeros((1, N[1], N[0], 2 if system == 0 else 4))
    U[0, ..., :nv2] = U2
    grid = acx.Grid(n=(N[0], N[1], 1), h=h, system=system, bc=(bc[0], bc[1], 0), gamma_hi=gamma[0],
                    gamma_lo=gamma[1])
    print(f"[case] 2d cweno4 system={system} bc={bc} gamma={gamma} {kind}", flush=True)
    R2, g2 = oracle_rhs2d(orc, N, h, system, bc, U2, gamma)
    got = acx.rhs(grid, torch.from_numpy(U).cuda()).cpu().numpy()
    err = rel_l2(got[0, ..., :nv2], R2)
    lin, _ = oracle_rhs2d(orc, N, h, system, bc, U2)
    print(f"[case] rhs rel_l2 {err:.3e} (linear scheme differs by {rel_l2(lin, R2):.2e})", flush=True)
    assert err <= TOL
    assert rel_l2(lin, R2) > 1e-6        # the weights did leave the linear weights
    dt = acx.cfl_dt(grid, 0.4)
    ref = U2.copy()
    for _ in range(3):
        orc.ts_solve("orc_rhs2d_cb", g2, ref, dt, dt, 0, 4)
    Ud = torch.from_numpy(U).cuda()
    for _ in range(3):
        acx.ssp_rks3_step(grid, Ud, dt)
    got = Ud.cpu().numpy()
    err = rel_l2(got[0, ..., :nv2], ref)
    print(f"[case] steps rel_l2 {err:.3e}", flush=True)
    assert err <= TOL
    if system == 1:
        assert np.all(got[..., 3] == 0.0)


def test_2d_cweno4_fp32_compensated(acx, torch_gpu):
    torch = torch_gpu
    N = 64
    grid = acx.Grid(n=(N, N, 1), h=1.0 / N, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    U = np.zeros(grid.shape())
    U[0, ..., :3] = _field2d(5, (N, N), 3, "disc") * 1e-3
    dt = acx.cfl_dt(grid, 0.4)
    ref = torch.from_numpy(U).cuda()
    st = acx.CompensatedState(grid, torch.from_numpy(U).cuda())
    for _ in range(20):
        acx.ssp_rks3_step(grid, ref, dt)
        st.step(dt)
    rel, mx = acx.diff_f32c(st.hi, st.lo, ref)
    print(f"[case] 2d cweno4 f32c rel_l2={rel:.2e} max={mx:.2e}", flush=True)
    assert rel < 1e-6


@pytest.mark.parametrize("N", [16, 32, 64])
def test_standing_wave_program_reproduces_error_dat(N):
    with open(os.path.join(ROOT, "tests", "golden", "standing_wave_2d_error.json")) as f:
        row = [r for r in json.load(f)["rows"] if r["N"] == N][0]
    exe = os.path.join(ROOT, "tests", "drivers", "standing_wave_2d")
    print(f"[case] {exe} {N}", flush=True)
    r = subprocess.run([exe, str(N)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    n, l2, li, steps = r.stdout.split()
    print(f"[case] N={N}: l2 {float(l2):.7e} ({row['l2_text']}), linf {float(li):.7e} ({row['linf_text']})", flush=True)
    assert int(n) == N and int(steps) == 2 * N
    assert abs(float(l2) / row["l2"] - 1) < 2e-7
    assert abs(float(li) / row["linf"] - 1) < 2e-7
