"""Batched ensemble: the persistent work-queue kernel must reproduce the
single-problem FP32 stepper bit for bit on every problem, through eager
steps and through replays of the captured hipGraph, for every tile
configuration the autotuner can pick; and each problem stays within FP32
accuracy of the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [(20, 16, 12), (16, 24, 20), (36, 12, 16), (12, 12, 28), (70, 9, 10), (24, 20, 16)]


def fields(acx, dims, h, seed):
    out = []
    for p, d in enumerate(dims):
        g = acx.Grid(n=d, h=h)
        U = acx.field_gaussians(g, seed + p, npulse=1 + p % 3, amp=(1e-4, 1e-3), sigma=(0.1, 0.2))
        out.append(U)
    return out


def single_steps(acx, torch, dims, h, U0s, dt, nsteps):
    res = []
    for d, U0 in zip(dims, U0s):
        g = acx.Grid(n=d, h=h)
        st = acx.CompensatedState(g, torch.from_numpy(U0).cuda())
        for _ in range(nsteps):
            st.step(dt)
        res.append((st.hi.cpu().numpy().ravel(), st.lo.cpu().numpy().ravel()))
    return res


def test_ensemble_equals_single_problem_stepper(acx, torch_gpu):
    torch = torch_gpu
    from acoustic_petsc_amd.ensemble import Ensemble
    h = 1.0 / 32
    dt = 0.4 * h / 3
    U0s = fields(acx, DIMS, h, 10)
    ref = single_steps(acx, torch, DIMS, h, U0s, dt, 3)
    for cfg in range(Ensemble.num_configs()):
        print(f"[case] ensemble eager config {cfg}", flush=True)
        e = Ensemble(DIMS, h)
        e.set_config(cfg)
        e.load(U0s)
        for _ in range(3):
            e.step(dt)
        hi, lo = e.state32()
        hi, lo = hi.cpu().numpy(), lo.cpu().numpy()
        for p in range(len(DIMS)):
            a, b = 4 * e.offsets[p], 4 * e.offsets[p + 1]
            assert np.array_equal(hi[a:b], ref[p][0]), (cfg, p)
            assert np.array_equal(lo[a:b], ref[p][1]), (cfg, p)
        e.close()


def test_graph_replays_match_eager(acx, torch_gpu):
    torch = torch_gpu
    from acoustic_petsc_amd.ensemble import Ensemble
    h = 1.0 / 32
    dt = 0.4 * h / 3
    U0s = fields(acx, DIMS, h, 20)
    eager = Ensemble(DIMS, h)
    graph = Ensemble(DIMS, h)
    eager.load(U0s)
    graph.load(U0s)
    for rep in range(3):
        print(f"[case] graph replay {rep}", flush=True)
        eager.step(dt)
        eager.step(dt)
        graph.graph_steps(dt, 2)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum()   # unrelated eager work between replays
        torch.cuda.synchronize()
        a, b = eager.state32(), graph.state32()
        assert np.array_equal(a[0].cpu().numpy(), b[0].cpu().numpy())
        assert np.array_equal(a[1].cpu().numpy(), b[1].cpu().numpy())
        del junk
    eager.close()
    graph.close()


def test_autotune_restores_state(acx, torch_gpu):
    torch = torch_gpu
    from acoustic_petsc_amd.ensemble import Ensemble
    h = 1.0 / 32
    dt = 0.4 * h / 3
    U0s = fields(acx, DIMS, h, 30)
    e = Ensemble(DIMS, h)
    e.load(U0s)
    before = e.state32()
    best, ms = e.autotune(dt)
    after = e.state32()
    print(f"[case] autotune picked {best} ({ms:.3f} ms/step)", flush=True)
    assert 0 <= best < e.num_configs() and e.config() == best and ms > 0
    assert np.array_equal(before[0].cpu().numpy(), after[0].cpu().numpy())
    e.close()


def test_ensemble_error_against_fp64_oracle(acx, orc, torch_gpu):
    torch = torch_gpu
    from acoustic_petsc_amd.ensemble import Ensemble
    h = 1.0 / 32
    dt = 0.4 * h / 3
    U0s = fields(acx, DIMS[:4], h, 40)
    e = Ensemble(DIMS[:4], h)
    e.load(U0s)
    for _ in range(2):
        e.step(dt)
    got = e.state64()
    for p, d in enumerate(DIMS[:4]):
        ref = U0s[p].copy()
        g = orc.grid3d(d, h, orc.ACOUSTIC4)
        orc.ts_solve("orc_rhs3d_cb_faithful", g, ref, dt, 2 * dt, 0, 4)
        err = np.linalg.norm(got[p] - ref) / np.linalg.norm(ref)
        print(f"[case] problem {p} {d}: rel L2 vs FP64 oracle {err:.2e}", flush=True)
        assert err < 1e-6
    e.close()
