"""GPU parity of the fused RHS / SSP stage kernels against the FP64 oracle.

Each case prints its name, shape and seed before the GPU call so that a
failing run names it.  Tolerance: relative L2 <= 1e-12 (FP64 kernels).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def make_field(acx, grid, seed, kind):
    rng = np.random.default_rng(seed)
    shape = grid.shape()
    if kind == "smooth":
        U = acx.field_gaussians(grid, seed, npulse=3, amp=(0.5, 1.0), sigma=(0.1, 0.3))
        for v in range(1, grid.nvar):
            U[..., v] = np.roll(U[..., 0], 3 * v, axis=v % 3) * (0.3 * v)
        return U
    if kind == "rough":
        return rng.standard_normal(shape)
    if kind == "step":
        nz, ny, nx, nv = shape
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        inside = ((x - nx / 2) ** 2 + 1.3 * (y - ny / 2) ** 2 + 0.7 * (z - nz / 2) ** 2) < (min(nx, ny) / 3) ** 2
        U = np.zeros(shape)
        for v in range(nv):
            U[..., v] = np.where(inside, 1.0 + 0.25 * v, 0.1) + 0.01 * rng.standard_normal(shape[:3])
        return U
    raise ValueError(kind)


def orc_grid(orc, grid):
    return orc.grid3d(tuple(grid.n), grid.h, grid.system, (grid.gamma_hi, grid.gamma_lo), grid.c0, grid.rho0,
                      0.4, tuple(grid.bc))


CASES = [
    # name, n, system, gamma, bc, kind
    ("ac4-lin-smooth", (70, 18, 12), 1, (1.0, 0.0), (0, 0, 0), "smooth"),
    ("ac4-lin-rough", (24, 20, 16), 1, (1.0, 0.0), (0, 0, 0), "rough"),
    ("ac4-nl-smooth", (24, 20, 16), 1, (0.85, 0.85), (0, 0, 0), "smooth"),
    ("ac4-nl-rough", (40, 30, 12), 1, (0.85, 0.85), (0, 0, 0), "rough"),
    ("ac4-nl-step", (32, 32, 10), 1, (0.85, 0.85), (0, 0, 0), "step"),
    ("ac4-nl-g05", (20, 16, 14), 1, (0.5, 0.3), (0, 0, 0), "rough"),
    ("w2-lin-per", (24, 20, 16), 0, (1.0, 0.0), (0, 0, 0), "rough"),
    ("w2-lin-adia", (24, 20, 16), 0, (1.0, 0.0), (1, 1, 1), "rough"),
    ("w2-lin-mixed", (70, 12, 10), 0, (1.0, 0.0), (0, 1, 1), "rough"),
    ("w2-nl-per", (24, 20, 16), 0, (0.85, 0.85), (0, 0, 0), "rough"),
    ("w2-nl-adia", (24, 20, 16), 0, (0.85, 0.85), (1, 1, 1), "rough"),
    # LinearisedEuler walls: 2 = reflective, 3 = transmissive
    ("ac4-lin-refl", (24, 20, 16), 1, (1.0, 0.0), (2, 2, 2), "rough"),
    ("ac4-lin-trans", (24, 20, 16), 1, (1.0, 0.0), (3, 3, 3), "rough"),
    ("ac4-lin-mixed", (70, 12, 10), 1, (1.0, 0.0), (2, 0, 3), "smooth"),
    ("ac4-nl-refl", (24, 20, 16), 1, (0.85, 0.85), (2, 2, 2), "rough"),
    ("ac4-nl-mixed", (20, 24, 14), 1, (0.85, 0.85), (3, 2, 0), "step"),
    # smallest grids the ABI accepts (4 cells per axis): stencils wrap / mirror more than once
    ("ac4-lin-tiny", (4, 5, 6), 1, (1.0, 0.0), (0, 2, 3), "rough"),
    ("ac4-nl-tiny", (5, 4, 4), 1, (0.85, 0.85), (0, 0, 2), "rough"),
    ("w2-nl-tiny", (4, 4, 5), 0, (0.85, 0.85), (1, 0, 1), "rough"),
    # all-periodic linear grids take the LDS-DMA marches: smaller than the
    # halo (every plane and row loaded several times) and ragged on every axis
    ("ac4-lin-tiny-per", (4, 5, 4), 1, (1.0, 0.0), (0, 0, 0), "rough"),
    ("w2-lin-tiny-per", (5, 4, 4), 0, (1.0, 0.0), (0, 0, 0), "rough"),
    ("w2-lin-per-ragged", (70, 33, 21), 0, (1.0, 0.0), (0, 0, 0), "smooth"),
]


@pytest.mark.parametrize("name,n,system,gamma,bc,kind", CASES, ids=[c[0] for c in CASES])
def test_rhs_parity(name, n, system, gamma, bc, kind, acx, orc, torch_gpu):
    torch = torch_gpu
    grid = acx.Grid(n=n, h=1.0 / n[0], system=system, bc=bc, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = make_field(acx, grid, 1234, kind)
    print(f"[case] rhs {name} n={n} seed=1234", flush=True)
    ref = orc.rhs3d(orc_grid(orc, grid), U, faithful=True)
    Ud = torch.from_numpy(U).cuda()
    out = acx.rhs(grid, Ud)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = rel_l2(got, ref)
    print(f"[case] rhs {name} rel_l2={err:.3e}", flush=True)
    assert np.isfinite(got).all()
    assert err <= TOL, f"{name}: rel L2 {err:.3e}"


@pytest.mark.parametrize("name,n,system,gamma,bc,kind", [CASES[0], CASES[1], CASES[3], CASES[7], CASES[13]],
                         ids=["ac4-lin-smooth", "ac4-lin-rough", "ac4-nl-rough", "w2-lin-adia", "ac4-lin-mixed"])
def test_forced_general_kernel_matches(name, n, system, gamma, bc, kind, acx, orc, torch_gpu, monkeypatch):
    """The general CWENO kernel run with the linear weights must agree with the oracle too."""
    torch = torch_gpu
    monkeypatch.setenv("ACX_FORCE_CWENO", "1")
    grid = acx.Grid(n=n, h=1.0 / n[0], system=system, bc=bc, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = make_field(acx, grid, 99, kind)
    print(f"[case] forced-cweno {name} n={n}", flush=True)
    ref = orc.rhs3d(orc_grid(orc, grid), U)
    got = acx.rhs(grid, torch.from_numpy(U).cuda()).cpu().numpy()
    err = rel_l2(got, ref)
    print(f"[case] forced-cweno {name} rel_l2={err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_rks3_step_parity(gamma, acx, orc, torch_gpu):
    torch = torch_gpu
    n = (24, 20, 18)
    grid = acx.Grid(n=n, h=1.0 / 24, system=acx.ACOUSTIC4, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = make_field(acx, grid, 7, "rough")
    dt = 0.4 * grid.h / 3.0
    print(f"[case] rks3 step gamma={gamma}", flush=True)
    ref = U.copy()
    g = orc_grid(orc, grid)
    for _ in range(3):
        orc.ts_solve("orc_rhs3d_cb_faithful", g, ref, dt, dt, 0, 4)
    Ud = torch.from_numpy(U).cuda()
    for _ in range(3):
        acx.ssp_rks3_step(grid, Ud, dt)
    got = Ud.cpu().numpy()
    err = rel_l2(got, ref)
    print(f"[case] rks3 gamma={gamma} rel_l2={err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("bc", [(2, 3, 0), (3, 0, 2)])
def test_walls_rks3_fp64_and_f32c(bc, acx, orc, torch_gpu):
    """Reflective / transmissive walls through full FP64 steps (oracle, 1e-12)
    and through the FP32 compensated stepper (FP32 accuracy)."""
    torch = torch_gpu
    n = (26, 20, 18)
    grid = acx.Grid(n=n, h=1.0 / 26, system=acx.ACOUSTIC4, bc=bc)
    U = make_field(acx, grid, 9, "smooth") * 1e-3
    dt = 0.4 * grid.h / 3.0
    print(f"[case] walls {bc}", flush=True)
    ref = U.copy()
    g = orc_grid(orc, grid)
    for _ in range(3):
        orc.ts_solve("orc_rhs3d_cb_faithful", g, ref, dt, dt, 0, 4)
    Ud = torch.from_numpy(U).cuda()
    for _ in range(3):
        acx.ssp_rks3_step(grid, Ud, dt)
    err = rel_l2(Ud.cpu().numpy(), ref)
    st = acx.CompensatedState(grid, torch.from_numpy(U).cuda())
    for _ in range(3):
        st.step(dt)
    err32 = rel_l2(st.value64().cpu().numpy(), ref)
    print(f"[case] walls {bc}: fp64 {err:.2e} f32c {err32:.2e}", flush=True)
    assert err <= TOL
    assert err32 < 1e-6


def test_stage_epilogues(acx, orc, torch_gpu):
    """OUT = a*U0 + b*Q + c*F(Q) for the three epilogue forms."""
    torch = torch_gpu
    n = (20, 16, 12)
    grid = acx.Grid(n=n, h=1.0 / 20, system=acx.ACOUSTIC4)
    rng = np.random.default_rng(5)
    Q = rng.standard_normal(grid.shape())
    U0 = rng.standard_normal(grid.shape())
    F = orc.rhs3d(orc_grid(orc, grid), Q)
    Qd, U0d = torch.from_numpy(Q).cuda(), torch.from_numpy(U0).cuda()
    for (a, b, c) in [(0.0, 0.0, 1.0), (0.0, 1.0, 0.25), (2.0 / 3.0, 1.0 / 3.0, 0.125)]:
        out = torch.empty_like(Qd)
        acx.stage(grid, Qd, U0d if a else None, out, a, b, c)
        ref = a * U0 + b * Q + c * F
        err = rel_l2(out.cpu().numpy(), ref)
        print(f"[case] stage a={a} b={b} c={c} rel_l2={err:.3e}", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("system,gamma,bc", [(0, (1.0, 0.0), (0, 0, 0)), (0, (1.0, 0.0), (1, 0, 1)),
                                             (1, (0.85, 0.85), (0, 0, 0)), (1, (1.0, 0.0), (0, 0, 0))],
                         ids=["w2-lin-per", "w2-lin-adia", "ac4-nl-per", "ac4-lin-per"])
def test_f32c_steps_track_fp64_oracle(system, gamma, bc, acx, orc, torch_gpu):
    """The FP32 compensated stepper (both systems, linear and nonlinear weights,
    walls) against the FP64 oracle after five steps, at FP32 accuracy."""
    torch = torch_gpu
    n = (28, 22, 18)
    grid = acx.Grid(n=n, h=1.0 / 28, system=system, bc=bc, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = make_field(acx, grid, 21, "smooth")
    dt = 0.4 * grid.h / 3.0
    print(f"[case] f32c system={system} gamma={gamma} bc={bc}", flush=True)
    ref = U.copy()
    g = orc_grid(orc, grid)
    for _ in range(5):
        orc.ts_solve("orc_rhs3d_cb_faithful", g, ref, dt, dt, 0, 4)
    st = acx.CompensatedState(grid, torch.from_numpy(U).cuda())
    for _ in range(5):
        st.step(dt)
    err = rel_l2(st.value64().cpu().numpy(), ref)
    print(f"[case] f32c system={system} gamma={gamma} bc={bc}: rel_l2 {err:.2e}", flush=True)
    assert err < 1e-6


def test_f32c_wave2_nonlinear_is_refused(acx, torch_gpu):
    torch = torch_gpu
    grid = acx.Grid(n=(16, 16, 16), h=1.0 / 16, system=acx.WAVE2, gamma_hi=0.85, gamma_lo=0.85)
    st = acx.CompensatedState(grid, torch.zeros(grid.shape(), dtype=torch.float64, device="cuda"))
    with pytest.raises(acx.AcxError, match="ACOUSTIC4"):
        st.step(1e-3)


@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_z_chunk_does_not_change_results(gamma, acx, torch_gpu, monkeypatch):
    """The z-chunk of a launch (chosen per launch from the occupancy model, or
    forced by ACX_ZC_LINEAR / ACX_ZC_CWENO) only changes which workgroup marches
    which planes: FP64 steps and FP32 compensated steps are bitwise equal for
    every chunk, including chunks that do not divide n_z."""
    torch = torch_gpu
    n = (40, 36, 44)
    grid = acx.Grid(n=n, h=1.0 / 40, system=acx.ACOUSTIC4, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U0 = torch.from_numpy(acx.field_gaussians(grid, 77, npulse=3, amp=(1e-4, 1e-3), sigma=(0.05, 0.1))).cuda()
    dt = acx.cfl_dt(grid, 0.4)
    knob = "ACX_ZC_LINEAR" if gamma == (1.0, 0.0) else "ACX_ZC_CWENO"
    results = []
    for zc in (None, 5, 16, 44):
        print(f"[case] z-chunk {knob}={zc} gamma={gamma}", flush=True)
        if zc is None:
            monkeypatch.delenv(knob, raising=False)
        else:
            monkeypatch.setenv(knob, str(zc))
        U = U0.clone()
        w = (torch.empty_like(U), torch.empty_like(U))
        for _ in range(2):
            acx.ssp_rks3_step(grid, U, dt, w)
        st = acx.CompensatedState(grid, U0)
        for _ in range(2):
            st.step(dt)
        results.append((U, st.hi.clone(), st.lo.clone()))
    for U, hi, lo in results[1:]:
        assert torch.equal(U, results[0][0])
        assert torch.equal(hi, results[0][1]) and torch.equal(lo, results[0][2])


def test_multi_round_launch_stores_rows_whole(acx, torch_gpu, monkeypatch):
    """FP64 linear stages of a launch with more workgroups than the GPU holds
    at once store each output row whole after a lane exchange; one that fits
    stores per cell.  Both give the same bits, here on a grid whose last
    column and row tiles are partial (90 = 2 x 32 + 26, 70 = 4 x 16 + 6):
    z-chunks of 2 and 3 planes make 480 and 330 workgroups, 64 makes 15."""
    torch = torch_gpu
    n = (90, 70, 64)
    grid = acx.Grid(n=n, h=1.0 / 90, system=acx.ACOUSTIC4)
    U0 = torch.from_numpy(acx.field_gaussians(grid, 909, npulse=3, amp=(1e-4, 1e-3), sigma=(0.05, 0.1))).cuda()
    dt = acx.cfl_dt(grid, 0.4)
    results = []
    for zc in (64, 2, 3, None):
        print(f"[case] row stores ACX_ZC_LINEAR={zc}", flush=True)
        if zc is None:
            monkeypatch.delenv("ACX_ZC_LINEAR", raising=False)
        else:
            monkeypatch.setenv("ACX_ZC_LINEAR", str(zc))
        U = U0.clone()
        w = (torch.empty_like(U), torch.empty_like(U))
        F = acx.rhs(grid, U)
        for _ in range(2):
            acx.ssp_rks3_step(grid, U, dt, w)
        results.append((F, U))
    for F, U in results[1:]:
        assert torch.equal(F, results[0][0]) and torch.equal(U, results[0][1])
