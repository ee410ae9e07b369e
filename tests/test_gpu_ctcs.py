"""CTCS leapfrog sibling (3d/ctcs/point_source/wave.c:109-240) on the GPU
through acx_ctcs_update_f64.

* The reference program (100^3, dx = 0.5, dt = 5e-4, c = 250, 4th-order
  Laplacian, source (50,50,50), receiver (57,57,57), 251 steps) reproduces
  the oracle's receiver trace to rounding and the reference's printed trace
  3d/ctcs/point_source/results/p.dat (tests/golden/ctcs_point_source_p.json)
  within the rounding of its %e text.
* Non-cubic grids, unequal spacings, both Laplacian orders and a per-cell
  wave speed field: the full field after several updates from random states
  matches the oracle.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu_steps(acx, torch, n, d, dt, stencil, src, rcv, Pold, P0, Cf, nt, it0=0, f0=100.0):
    import ctypes as C
    cs = acx._Ctcs()
    cs.n[:] = list(n)
    cs.stencil = stencil
    cs.src[:] = list(src)
    cs.rcv[:] = list(rcv)
    cs.dx, cs.dy, cs.dz, cs.dt, cs.f0 = d[0], d[1], d[2], dt, f0
    a = torch.from_numpy(np.ascontiguousarray(Pold)).cuda()
    b = torch.from_numpy(np.ascontiguousarray(P0)).cuda()
    new = torch.zeros_like(b)
    cf = torch.from_numpy(np.ascontiguousarray(Cf)).cuda()
    r = (C.c_double * 1)()
    tr = []
    for it in range(nt):
        acx._check(acx.lib().acx_ctcs_update_f64(C.byref(cs), a.data_ptr(), b.data_ptr(), new.data_ptr(),
                                                 cf.data_ptr(), it0 + it, r, acx._stream()))
        tr.append(r[0])
        a, b, new = b, new, a
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy(), np.array(tr)


def test_reference_program_trace(acx, orc, torch_gpu):
    p = acx.Ctcs()
    print(f"[case] ctcs reference program n={p.n} stencil={p.stencil} nt=251", flush=True)
    tr = np.array(acx.ctcs_run(p, 251))
    ref = orc.ctcs_trace()
    scale = np.max(np.abs(ref))
    assert scale > 0
    assert np.max(np.abs(tr - ref)) < 1e-12 * scale
    with open(os.path.join(ROOT, "tests", "golden", "ctcs_point_source_p.json")) as f:
        g = json.load(f)
    texts = g["p_text"]
    assert len(texts) == 251
    for i, (v, t) in enumerate(zip(tr, texts)):
        # "%e" keeps 7 significant digits: the printed value is within half a
        # unit of the 6th decimal of the mantissa; the late tail (|p| ~ 1e-13,
        # 1e-8 of the peak) is a cancellation residue, where rounding at the
        # 1e-12-of-peak level (FMA contraction, another summation order) can
        # move the 7th digit
        half = 0.5 * 10.0 ** (int(t.split("e")[1]) - 6)
        assert abs(v - float(t)) <= half * (1 + 1e-9) + 1e-12 * scale, (i, v, t)


# even x extents take the LDS-DMA march (including a grid smaller than its
# halo and one with partial tiles on every axis), odd ones the register march
@pytest.mark.parametrize("n,d,stencil", [((20, 18, 16), (0.5, 0.4, 0.3), 1),
                                         ((20, 18, 16), (0.5, 0.4, 0.3), 2),
                                         ((6, 5, 5), (0.3, 0.3, 0.3), 2),
                                         ((34, 20, 21), (0.2, 0.3, 0.25), 2),
                                         ((7, 33, 9), (0.2, 0.25, 0.2), 2)])
def test_fields_match_oracle(acx, orc, torch_gpu, n, d, stencil):
    rng = np.random.default_rng(7 + stencil)
    shape = (n[2], n[1], n[0])
    Pold = rng.standard_normal(shape)
    P0 = rng.standard_normal(shape)
    Cf = 0.5 + rng.random(shape)
    dt = 0.2 * min(d) / Cf.max()
    src, rcv = (n[0] // 2, n[1] // 3, n[2] - 1), (1, n[1] - 1, 0)
    print(f"[case] ctcs fields n={n} d={d} stencil={stencil}", flush=True)
    oa, ob, otr = orc.ctcs_steps(n, d, dt, stencil, src, rcv, Pold, P0, Cf, 12, it0=3)
    ga, gb, gtr = _gpu_steps(acx, torch_gpu, n, d, dt, stencil, src, rcv, Pold, P0, Cf, 12, it0=3)
    for o, g in ((oa, ga), (ob, gb)):
        assert np.linalg.norm(g - o) / np.linalg.norm(o) < 1e-13
        assert np.max(np.abs(g - o)) < 1e-12 * np.max(np.abs(o))
    assert np.max(np.abs(gtr - otr)) < 1e-12 * np.max(np.abs(otr))


@pytest.mark.parametrize("n", [(64, 64, 1), (128, 1, 1)], ids=["2d-plane-wave", "1d-travelling-wave"])
def test_2d_1d_programs(acx, orc, torch_gpu, n):
    """The 2D and 1D CTCS programs (2d/ctcs/plane_wave/wave.c:28-117,161-202,
    1d/ctcs/travelling_wave/wave.c) as grids with axes of one cell:
    p = sin(2 pi (x - t)) sampled at the nodes x = i h, dt = 0.5 CFL h / c with
    CFL 0.5, two exact start levels; the fields match the oracle and the
    error against the exact wave falls at second order."""
    errs = []
    for N in (n[0] // 2, n[0]):
        nn = (N, 1 if n[1] == 1 else N, 1)
        h = 1.0 / N
        dt = 0.5 * 0.5 * h
        x = np.arange(N) * h
        shape = (1, nn[1], N)
        P0 = np.broadcast_to(np.sin(2 * np.pi * x), shape).copy()
        P1 = np.broadcast_to(np.sin(2 * np.pi * (x - dt)), shape).copy()
        Cf = np.ones(shape)
        nt = int(round(0.25 / dt))
        d = (h, h if nn[1] > 1 else 1.0, 1.0)
        print(f"[case] ctcs {nn} nt={nt}", flush=True)
        oa, ob, _ = orc.ctcs_steps(nn, d, dt, 1, (-1, -1, -1), (0, 0, 0), P0, P1, Cf, nt)
        ga, gb, _ = _gpu_steps(acx, torch_gpu, nn, d, dt, 1, (-1, -1, -1), (0, 0, 0), P0, P1, Cf, nt)
        assert np.max(np.abs(gb - ob)) < 1e-12
        exact = np.sin(2 * np.pi * (x - (nt + 1) * dt))
        errs.append(np.max(np.abs(gb - exact)))
    rate = np.log2(errs[0] / errs[1])
    print(f"[case] ctcs {n}: errors {errs[0]:.3e} {errs[1]:.3e} rate {rate:.2f}", flush=True)
    assert 1.8 < rate < 2.2
