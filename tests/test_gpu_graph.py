"""The FP64 stepper replayed from a captured hipGraph (the four stage launches
of one step, acx_ssp_rks3_steps_graph_f64) against the eager stepper: bit for
bit, over several replays with unrelated eager work between them, on 3D and
2D grids, linear and nonlinear weights, and again after dt or the arrays
change (a new capture)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [((40, 36, 28), (1.0, 0.0)), ((24, 20, 16), (0.85, 0.85)), ((64, 48, 1), (1.0, 0.0)),
         ((48, 40, 1), (0.85, 0.85))]


@pytest.mark.parametrize("n,gamma", CASES, ids=["3d-lin", "3d-nl", "2d-lin", "2d-nl"])
def test_graph_steps_equal_eager_steps(acx, torch_gpu, n, gamma):
    torch = torch_gpu
    g = acx.Grid(n=n, h=1.0 / n[0], gamma_hi=gamma[0], gamma_lo=gamma[1])
    U0 = acx.field_gaussians(g, 606, npulse=3, amp=(1e-4, 1e-3), sigma=(0.05, 0.12))
    if n[2] == 1:
        U0[..., 3] = 0.0
    dt = acx.cfl_dt(g, 0.4)
    print(f"[case] graph steps n={n} gamma={gamma}", flush=True)
    a = torch.from_numpy(U0).cuda()
    b = a.clone()
    wa = (torch.empty_like(a), torch.empty_like(a))
    wb = (torch.empty_like(b), torch.empty_like(b))
    for _ in range(5):
        acx.ssp_rks3_step(g, a, dt, wa)
    acx.ssp_rks3_steps_graph(g, b, dt, 2, wb)
    junk = torch.randn(1 << 20, device="cuda")           # unrelated eager work between replays
    junk.mul_(2.0).add_(1.0)
    torch.cuda.synchronize()
    acx.ssp_rks3_steps_graph(g, b, dt, 1, wb)
    float(junk.sum())
    acx.ssp_rks3_steps_graph(g, b, dt, 2, wb)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_graph_recaptures_when_dt_or_arrays_change(acx, torch_gpu):
    torch = torch_gpu
    g = acx.Grid(n=(32, 24, 20), h=1.0 / 32)
    U0 = torch.from_numpy(acx.field_gaussians(g, 7, npulse=2)).cuda()
    dt = acx.cfl_dt(g, 0.4)
    a, b, c = U0.clone(), U0.clone(), U0.clone()
    w = (torch.empty_like(a), torch.empty_like(a))
    w2 = (torch.empty_like(a), torch.empty_like(a))
    acx.ssp_rks3_steps_graph(g, b, dt, 3, w)
    acx.ssp_rks3_steps_graph(g, b, 0.5 * dt, 2, w)       # new dt: new capture
    acx.ssp_rks3_steps_graph(g, c, dt, 3, w2)             # new arrays: new capture
    for _ in range(3):
        acx.ssp_rks3_step(g, a, dt, w)
    assert torch.equal(a, c)
    for _ in range(2):
        acx.ssp_rks3_step(g, a, 0.5 * dt, w)
    assert torch.equal(a, b)
