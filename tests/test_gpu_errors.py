"""Non-finite states and the FP64 reductions on the GPU.

The reference aborts on an inadmissible state (MPI_Abort in PDEConsFlux,
3d/weno3/LinearisedEuler/pde.c:50-55).  Here every stage kernel raises a
per-device flag when it writes a NaN or Inf: acx_nonfinite_check reports it
after synchronising, every later RHS / stage / step call refuses to run with
ACX_ENONFINITE (code 4), and petsclite's TSSolve returns the error instead of
a blown-up state.

The reductions are VecNorm (3d/weno4/plane_wave/error_norms.c:85-91) and the
volume-weighted energy sum (3d/weno4/Navier-Stokes-3D/kinetic_energy.c:24-39),
checked against numpy FP64 sums of the same arrays; a NaN propagates into the
norms as it does in PETSc.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "tests", "drivers")


@pytest.fixture
def clean(acx, torch_gpu):
    acx.nonfinite_clear()
    yield
    torch_gpu.cuda.synchronize()
    acx.nonfinite_clear()


def _state(acx, torch, n=24, gamma=(1.0, 0.0), seed=11):
    g = acx.Grid(n=(n, n, n), h=1.0 / n, system=acx.ACOUSTIC4, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = acx.field_gaussians(g, seed, npulse=3, amp=(1e-4, 1e-3), sigma=(0.05, 0.1))
    return g, torch.from_numpy(U).cuda()


def _expect_nonfinite(acx, fn):
    with pytest.raises(acx.AcxError) as ei:
        fn()
    assert ei.value.code == 4, ei.value


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_fp64_step_reports_a_nonfinite_state(acx, torch_gpu, clean, bad, gamma):
    torch = torch_gpu
    print(f"[case] fp64 step bad={bad} gamma={gamma}", flush=True)
    g, U = _state(acx, torch, gamma=gamma)
    dt = acx.cfl_dt(g, 0.4)
    work = (torch.empty_like(U), torch.empty_like(U))
    acx.ssp_rks3_step(g, U, dt, work)        # finite: no report
    acx.nonfinite_check()
    U[7, 5, 3, 0] = bad
    acx.ssp_rks3_step(g, U, dt, work)        # launched before the flag is raised
    _expect_nonfinite(acx, acx.nonfinite_check)
    _expect_nonfinite(acx, lambda: acx.ssp_rks3_step(g, U, dt, work))
    _expect_nonfinite(acx, lambda: acx.rhs(g, U))
    acx.nonfinite_clear()
    g2, U2 = _state(acx, torch, gamma=gamma, seed=12)
    acx.ssp_rks3_step(g2, U2, dt, work)
    acx.nonfinite_check()
    assert torch.isfinite(U2).all()


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_rhs_reports_a_nonfinite_input(acx, torch_gpu, clean, bad):
    torch = torch_gpu
    print(f"[case] rhs bad={bad}", flush=True)
    g, U = _state(acx, torch)
    U[3, 4, 5, 2] = bad
    acx.rhs(g, U)
    _expect_nonfinite(acx, acx.nonfinite_check)


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_compensated_fp32_step_reports_a_nonfinite_state(acx, torch_gpu, clean, bad):
    torch = torch_gpu
    print(f"[case] f32c step bad={bad}", flush=True)
    g, U = _state(acx, torch)
    st = acx.CompensatedState(g, U)
    dt = acx.cfl_dt(g, 0.4)
    st.step(dt)
    acx.nonfinite_check()
    st.hi[2, 9, 11, 1] = bad
    st.step(dt)
    _expect_nonfinite(acx, acx.nonfinite_check)
    _expect_nonfinite(acx, lambda: st.step(dt))


def test_large_grid_without_blowup_never_reports(acx, torch_gpu, clean):
    """The watch costs no false alarm on a long finite run at 128^3."""
    torch = torch_gpu
    g, U = _state(acx, torch, n=128)
    dt = acx.cfl_dt(g, 0.4)
    work = (torch.empty_like(U), torch.empty_like(U))
    for _ in range(20):
        acx.ssp_rks3_step(g, U, dt, work)
    acx.nonfinite_check()


@pytest.mark.parametrize("bad", ["nan", "inf"])
def test_tssolve_fails_on_a_nonfinite_state(tmp_path, bad):
    """The reference driver's TSSolve through petsclite: a NaN / Inf in the
    initial pressure makes TSSolve return the non-finite error (76 + 4)."""
    args = [os.path.join(DRV, "acoustic_ts"), "24", "3", "1.0", "0.0", "fused", str(tmp_path / "o.bin"), bad]
    print("[case]", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "error 80" in r.stderr and "NaN or Inf" in r.stderr, r.stderr


def test_tssolve_finite_run_succeeds(tmp_path):
    args = [os.path.join(DRV, "acoustic_ts"), "24", "3", "1.0", "0.0", "generic", str(tmp_path / "o.bin")]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


# ---- reductions (a11) against numpy FP64 ------------------------------------------

@pytest.mark.parametrize("comp", [-1, 0, 1, 2, 3])
def test_norms_match_numpy(acx, torch_gpu, comp):
    torch = torch_gpu
    rng = np.random.default_rng(40 + comp)
    X = rng.standard_normal((20, 18, 22, 4)) * np.exp(rng.uniform(-3, 3, (20, 18, 22, 4)))
    s, m = acx.norms(torch.from_numpy(X).cuda(), comp=comp)
    sel = X if comp < 0 else X[..., comp]
    want_s = float(np.sum(sel.astype(np.float64) ** 2))
    want_m = float(np.max(np.abs(sel)))
    print(f"[case] norms comp={comp}: {s:.17g} vs {want_s:.17g}", flush=True)
    assert abs(s / want_s - 1) < 1e-13
    assert m == want_m


def test_norms_propagate_nan_and_inf(acx, torch_gpu):
    torch = torch_gpu
    X = np.ones((8, 8, 8, 4))
    X[3, 2, 1, 2] = np.nan
    s, m = acx.norms(torch.from_numpy(X).cuda())
    assert np.isnan(s) and np.isnan(m)
    X[3, 2, 1, 2] = -np.inf
    s, m = acx.norms(torch.from_numpy(X).cuda())
    assert s == np.inf and m == np.inf
    # a component filter skips the other components' NaN
    X[3, 2, 1, 2] = np.nan
    s, m = acx.norms(torch.from_numpy(X).cuda(), comp=0)
    assert s == 8 ** 3 and m == 1.0


@pytest.mark.parametrize("n,rho0,c0", [((20, 16, 12), 1.0, 1.0), ((18, 22, 1), 1.3, 0.7)])
def test_energy_matches_numpy(acx, torch_gpu, n, rho0, c0):
    """E = 1/2 sum (p^2 / K + rho0 |u|^2) h^3 (h^2 on a one-plane 2D grid), K = rho0 c0^2."""
    torch = torch_gpu
    g = acx.Grid(n=n, h=0.05, system=acx.ACOUSTIC4, rho0=rho0, c0=c0)
    rng = np.random.default_rng(3)
    U = rng.standard_normal(g.shape())
    if n[2] == 1:
        U[..., 3] = 0.0
    e = acx.energy(g, torch.from_numpy(U).cuda())
    K = rho0 * c0 * c0
    vol = g.h ** (2 if n[2] == 1 else 3)
    want = 0.5 * float(np.sum(U[..., 0] ** 2 / K + rho0 * np.sum(U[..., 1:] ** 2, axis=-1))) * vol
    print(f"[case] energy n={n}: {e:.17g} vs {want:.17g}", flush=True)
    assert abs(e / want - 1) < 1e-13


def test_diff_matches_numpy(acx, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(5)
    B = rng.standard_normal(30000)
    hi = B.astype(np.float32)
    lo = (B - hi.astype(np.float64)).astype(np.float32) + np.float32(1e-9) * rng.standard_normal(30000).astype(np.float32)
    rel, mx = acx.diff_f32c(torch.from_numpy(hi).cuda(), torch.from_numpy(lo).cuda(), torch.from_numpy(B).cuda())
    d = hi.astype(np.float64) + lo.astype(np.float64) - B
    assert abs(rel / (np.linalg.norm(d) / np.linalg.norm(B)) - 1) < 1e-12
    assert mx == float(np.max(np.abs(d)))


def test_reductions_on_two_streams(acx, torch_gpu):
    """Each call takes its own partials: norms issued on two streams give
    the values of the same calls issued alone."""
    torch = torch_gpu
    rng = np.random.default_rng(8)
    A = torch.from_numpy(rng.standard_normal((64, 64, 64, 4))).cuda()
    B = torch.from_numpy(3.0 * rng.standard_normal((64, 64, 64, 4))).cuda()
    ref_a, ref_b = acx.norms(A), acx.norms(B)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    import threading
    out = {}

    def run(name, X, s):
        for _ in range(10):
            out[name] = acx.norms(X, stream=s)

    ta = threading.Thread(target=run, args=("a", A, s1))
    tb = threading.Thread(target=run, args=("b", B, s2))
    ta.start(); tb.start(); ta.join(); tb.join()
    assert out["a"] == ref_a and out["b"] == ref_b


@pytest.mark.parametrize("n,off_x,off_y", [(1, 0, 0), (7, 0, 0), (1 << 20, 0, 0), (1 << 20, 1, 0),
                                           (12345, 0, 1), (4097, 1, 1), (262147, 0, 0)])
def test_vec_copy_matches(acx, torch_gpu, n, off_x, off_y):
    """VecCopy: bit-exact for even and odd sizes and for arrays 8 bytes off
    16-byte alignment (the scalar form); the bytes after Y's end stay."""
    torch = torch_gpu
    print(f"[case] vec_copy n={n} off=({off_x},{off_y})", flush=True)
    rng = np.random.default_rng(n)
    src = torch.from_numpy(rng.standard_normal(n + 2)).cuda()
    dst = torch.full((n + 3,), -7.0, dtype=torch.float64, device="cuda")
    X, Y = src[off_x:off_x + n], dst[off_y:off_y + n]
    acx.vec_copy(X, Y)
    d = dst.cpu().numpy()
    assert np.array_equal(d[off_y:off_y + n], src.cpu().numpy()[off_x:off_x + n])
    assert np.all(d[:off_y] == -7.0) and np.all(d[off_y + n:] == -7.0)


def test_vec_copy_refuses_overlap(acx, torch_gpu):
    torch = torch_gpu
    a = torch.zeros(100, dtype=torch.float64, device="cuda")
    with pytest.raises(acx.AcxError):
        acx.vec_copy(a[:50], a[10:60])
    acx.vec_copy(a, a)            # same array: nothing to do
