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
* Nonlinear weights are refused on 2D grids (the 3D nonlinear scheme on one
  plane is not the 2D CWENO4 of 2d/weno4/bubble).
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


def oracle_rhs2d(orc, N, h, system, bc, U2):
    g2 = orc.G2((N[0], N[1]), h, system, 0, 1.0, 0.0, 1.0, 1.0, bc)
    R = np.zeros_like(U2)
    assert orc.lib().orc_rhs2d_faithful(C.byref(g2), orc.P(U2), orc.P(R)) == 0
    return R, g2


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


def test_nonlinear_weights_refused_on_2d_grid(acx, torch_gpu):
    torch = torch_gpu
    grid = acx.Grid(n=(16, 16, 1), h=1.0 / 16, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    U = torch.zeros(grid.shape(), dtype=torch.float64, device="cuda")
    with pytest.raises(acx.AcxError, match="2D grid"):
        acx.rhs(grid, U)


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
