"""Learned troubled-cell indicator: MFMA forward and fp32 backward against a
CPU FP64 MLP on the same bf16-rounded operands, labels against the oracle's
CWENO4 weights, fused inference against features + forward, the selective
stepper's limits, and what a trained network gains over the method without it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def split(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[:1024].reshape(32, 32), p[1024:1056], p[1056:2080].reshape(32, 32), p[2080:2112], p[2112:2144],
            p[2144])


ACTS = {"relu": lambda z: np.maximum(z, 0.0), "tanh": np.tanh}


def cpu_forward(p, X, act="relu"):
    W1, b1, W2, b2, w3, b3 = split(p)
    W1r, W2r = bf16(W1).astype(np.float64), bf16(W2).astype(np.float64)
    x = bf16(X).astype(np.float64)
    f = ACTS[act]
    h1 = f(x @ W1r.T + b1)
    h1r = bf16(h1).astype(np.float64)
    h2 = f(h1r @ W2r.T + b2)
    y = 1.0 / (1.0 + np.exp(-(h2 @ w3 + b3)))
    return y, h1r, h2


@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_mfma_forward_matches_cpu(acx, torch_gpu, act):
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    rng = np.random.default_rng(0)
    p = learn.init_params(3, act)
    p[1024:1056] = rng.uniform(-0.2, 0.2, 32)
    p[2080:2112] = rng.uniform(-0.2, 0.2, 32)
    p[2144] = 0.1
    n = 1000   # not a multiple of 16
    X = np.zeros((n, 32), np.float32)
    X[:, :19] = rng.uniform(-1.5, 1.5, (n, 19))
    print(f"[case] mfma forward n=1000 {act}", flush=True)
    pd = torch.from_numpy(p).cuda()
    packed = learn.pack(pd, act=act)
    Y, H1, H2 = learn.forward(packed, torch.from_numpy(X).cuda(), keep=True)
    Y, H1, H2 = (a.cpu().numpy().astype(np.float64) for a in (Y, H1, H2))
    y, h1r, h2 = cpu_forward(p, X, act)
    # layer 1 (kept as bf16): one bf16 step apart at most, where the fp32 sum
    # or the activation's last bits land on the other side of a rounding tie
    d1 = np.abs(H1 - h1r)
    assert np.all(d1 <= 2.0 ** -7 * np.abs(h1r) + 1e-30)
    assert np.mean(d1 > 0) < 0.01
    # layer 2 and the output on the GPU's own layer-1 values: fp32 accuracy
    W1, b1, W2, b2, w3, b3 = split(p)
    h2g = ACTS[act](H1 @ bf16(W2).astype(np.float64).T + b2)
    assert np.max(np.abs(H2 - h2g)) < 2e-5
    assert np.max(np.abs(Y - 1.0 / (1.0 + np.exp(-(H2 @ w3 + b3))))) < 2e-6
    # the whole chain: a layer-1 rounding step moves y by < 1e-3
    assert np.max(np.abs(Y - y)) < 1e-3 and np.median(np.abs(Y - y)) < 1e-6


@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_backward_matches_cpu(acx, torch_gpu, act):
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    rng = np.random.default_rng(1)
    p = learn.init_params(5, act)
    n = 777
    X = np.zeros((n, 32), np.float32)
    X[:, :19] = rng.uniform(-1, 1, (n, 19))
    T = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    pd = torch.from_numpy(p).cuda()
    packed = learn.pack(pd, act=act)
    Xd = torch.from_numpy(X).cuda()
    Y, H1, H2 = learn.forward(packed, Xd, keep=True)
    pw = 3.0
    G = learn.backward(pd, Xd, H1, H2, Y, torch.from_numpy(T).cuda(), pw, act=act).cpu().numpy().astype(np.float64)
    # CPU FP64 gradient on the GPU's forward quantities
    y, h1r, h2 = Y.cpu().numpy().astype(np.float64), H1.cpu().numpy().astype(np.float64), H2.cpu().numpy().astype(np.float64)
    t = T.astype(np.float64)
    c = np.where(T > 0, pw, 1.0)
    ref = cpu_grad(p, bf16(X).astype(np.float64), bf16(split(p)[2]).astype(np.float64), y, h1r, h2, c, t, act)
    err = np.linalg.norm(G - ref) / np.linalg.norm(ref)
    print(f"[case] backward {act} rel err {err:.2e}", flush=True)
    assert err < 1e-5
    # the gradient formula itself against central differences of the FP64
    # loss (FP64 network, no bf16 rounding), for weights of every layer
    W1, b1, W2, b2, w3, b3 = split(p)
    x = X.astype(np.float64)

    f = ACTS[act]

    def fwd(pp):
        W1, b1, W2, b2, w3, b3 = split(pp)
        h1 = f(x @ W1.T + b1)
        h2 = f(h1 @ W2.T + b2)
        return 1.0 / (1.0 + np.exp(-(h2 @ w3 + b3))), h1, h2

    def loss(pp):
        yy = np.clip(fwd(pp)[0], 1e-12, 1 - 1e-12)
        return -np.sum(c * (t * np.log(yy) + (1 - t) * np.log(1 - yy))) / n
    y64, h164, h264 = fwd(p.astype(np.float64))
    g64 = cpu_grad(p, x, W2, y64, h164, h264, c, t, act)
    for k in (2144, 2120, 2090, 1100, 1060, 1030, 5, 300):
        e = 1e-6
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[k] += e
        pm[k] -= e
        fd = (loss(pp) - loss(pm)) / (2 * e)
        assert abs(fd - g64[k]) < 1e-5 * abs(fd) + 1e-6 * np.abs(g64).max(), (k, fd, g64[k])


def cpu_grad(p, x, W2, y, h1, h2, c, t, act="relu"):
    """Gradient of the class-weighted cross entropy of the MLP, from its
    forward quantities (inputs x, layer-2 weights W2 as the forward used them;
    the activation's derivative from its output: ReLU h > 0, tanh 1 - h^2)."""
    n = x.shape[0]
    w3 = split(p)[4]
    dz = c * (y - t) / n
    da = (lambda h: (h > 0).astype(np.float64)) if act == "relu" else (lambda h: 1.0 - h * h)
    dh2 = dz[:, None] * w3[None, :] * da(h2)
    dh1 = (dh2 @ W2) * da(h1)
    return np.concatenate([(dh1.T @ x).ravel(), dh1.sum(0), (dh2.T @ h1).ravel(), dh2.sum(0), (dz[:, None] * h2).sum(0),
                           [dz.sum()]])


def test_labels_match_oracle_weights(acx, orc, torch_gpu):
    """Label = L1 distance of the eight normalised CWENO weights from the
    linear weights, against the oracle's weights of the same 5^3 blocks."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    N = 12
    grid = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    rng = np.random.default_rng(2)
    U = rng.standard_normal(grid.shape()) * 1e-3
    U[..., 0] += np.where(np.arange(N)[None, None, :] < N // 2, 1e-3, 0.0)
    T, W = learn.labels(grid, torch.from_numpy(U).cuda(), thr=0.05)
    W = W.cpu().numpy().reshape(N, N, N, 4)
    T = T.cpu().numpy().reshape(N, N, N, 4)
    g4 = 0.85
    gam = np.array([g4, (1 - g4) * 0.85] + [(1 - g4) * (1 - 0.85) / 6] * 6)
    for (k, j, i, v) in [(0, 0, 0, 0), (3, 5, 5, 1), (6, 6, 6, 2), (11, 2, 9, 3), (4, 7, 1, 0)]:
        blk = np.take(np.take(np.take(U[..., v], range(k - 2, k + 3), axis=0, mode="wrap"),
                              range(j - 2, j + 3), axis=1, mode="wrap"), range(i - 2, i + 3), axis=2, mode="wrap")
        _, dg = orc.cweno3d(blk, (0.85, 0.85))
        dev = np.abs(np.array(dg[1:9]) - gam).sum()
        assert abs(W[k, j, i, v] - dev) < 1e-6 * max(dev, 1e-3)
        assert T[k, j, i, v] == (dev > 0.05)
    assert 0 < T.mean() < 1


@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_fused_indicator_equals_features_plus_forward(acx, torch_gpu, act):
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    N = 16
    grid = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    U = torch.from_numpy(learn.steep_pulse_field(grid, 4)).cuda().to(torch.float32)
    p = torch.from_numpy(learn.init_params(9, act)).cuda()
    packed = learn.pack(p, act=act)
    X = learn.features(grid, U)
    Y = learn.forward(packed, X).cpu().numpy().reshape(-1, 4)
    st = learn.SelectiveStepper(grid, packed, U.to(torch.float64), thr=0.5, dilate=0, judge=False)
    st.refresh_flags()
    fl = st.flags.cpu().numpy()
    want = ((Y > 0.5).astype(np.uint8) * np.array([1, 2, 4, 8], np.uint8)).sum(1)
    near = np.any(np.abs(Y - 0.5) < 1e-6, axis=1)
    assert np.all((fl == want) | near)


def test_flags_dilate(acx, torch_gpu):
    """Dilation: OR over the cell and its 6 periodic face neighbours, spread
    to all four component bits."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    import ctypes as C
    g = acx.Grid(n=(9, 7, 5), h=0.1)
    rng = np.random.default_rng(8)
    fl = np.zeros(g.ncells, np.uint8)
    idx = rng.choice(g.ncells, 6, replace=False)
    fl[idx] = rng.integers(1, 16, idx.size)
    src = torch.from_numpy(fl).cuda()
    dst = torch.empty_like(src)
    L = learn._bind()
    _, gp = learn._g(g)
    acx._check(L.acx_flags_dilate(gp, C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), acx._stream()))
    a = fl.reshape(5, 7, 9) != 0
    want = a.copy()
    for ax in range(3):
        want |= np.roll(a, 1, axis=ax) | np.roll(a, -1, axis=ax)
    assert np.array_equal(dst.cpu().numpy().reshape(5, 7, 9), np.where(want, 15, 0).astype(np.uint8))


def test_selective_limits(acx, torch_gpu):
    """All flags set == full CWENO4 FP32 step, no flag set == the linear scheme
    (both to FP32 rounding: the selective kernel is another instantiation)."""
    torch = torch_gpu
    N = 24
    gnl = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    glin = acx.Grid(n=(N, N, N), h=1.0 / N)
    from acoustic_petsc_amd import learn
    U = torch.from_numpy(learn.steep_pulse_field(gnl, 11)).cuda()
    dt = acx.cfl_dt(gnl, 0.4)
    ones = torch.full((gnl.ncells,), 15, dtype=torch.uint8, device="cuda")
    zeros = torch.zeros_like(ones)
    a, b = acx.CompensatedState(gnl, U), acx.CompensatedState(gnl, U)
    c, d = acx.CompensatedState(gnl, U), acx.CompensatedState(glin, U)
    for _ in range(3):
        a.step(dt, flags=ones)
        b.step(dt)   # general kernel, no flags
        c.step(dt, flags=zeros)
        d.step(dt)   # linear kernel
    ra, rb = a.value64().cpu().numpy(), b.value64().cpu().numpy()
    rc, rd = c.value64().cpu().numpy(), d.value64().cpu().numpy()
    # same arithmetic, different instantiations: equal to FP32 rounding
    assert np.linalg.norm(ra - rb) / np.linalg.norm(rb) < 1e-6
    assert np.linalg.norm(rc - rd) / np.linalg.norm(rd) < 1e-6
    assert np.linalg.norm(ra - rc) / np.linalg.norm(rc) > 1e-6   # the fronts do trigger the weights


def test_trained_indicator_gain(acx, torch_gpu):
    """Train on solver-labelled 64^3 snapshots (other seeds), then run a
    held-out steep pulse at 128^3: the learned stepper must skip most of the
    WENO work (the method without the indicator evaluates it everywhere),
    stay at the FP32 rounding level of the all-WENO FP64 solution, and far
    below the error of the linear scheme."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    X, T = learn.make_dataset(n_snapshots=4, N=64, seed0=300)
    print(f"[case] dataset {X.shape[0]} samples, positive {float(T.float().mean()):.3f}", flush=True)
    p = learn.train(X, T, epochs=4, batch=8192, lr=3e-3, seed=1)
    packed = learn.pack(p)
    rec, prec, frac = learn.evaluate(packed, X, T, 0.3)
    print(f"[case] train recall {rec:.3f} precision {prec:.3f} flagged {frac:.3f}", flush=True)
    assert rec > 0.95
    del X, T
    N = 128
    g = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    glin = acx.Grid(n=(N, N, N), h=1.0 / N)
    U0 = torch.from_numpy(learn.steep_pulse_field(g, 999)).cuda()
    dt = acx.cfl_dt(g, 0.4)
    ref = U0.clone()
    w = (torch.empty_like(ref), torch.empty_like(ref))
    sel = learn.SelectiveStepper(g, packed, U0)
    lin = acx.CompensatedState(glin, U0)
    fracs = []
    for _ in range(12):
        acx.ssp_rks3_step(g, ref, dt, w)
        sel.step(dt)
        lin.step(dt)
        fracs.append(sel.flagged_fraction())
    e_sel = acx.diff_f32c(sel.state.hi, sel.state.lo, ref)[0]
    e_lin = acx.diff_f32c(lin.hi, lin.lo, ref)[0]
    print(f"[case] flagged {np.mean(fracs):.3f}, rel L2 vs all-WENO FP64: learned {e_sel:.2e}, linear {e_lin:.2e}",
          flush=True)
    assert np.mean(fracs) < 0.15
    assert e_sel < 1e-6
    assert e_sel < 1e-3 * e_lin


@pytest.mark.parametrize("walls", [False, True])
def test_tile_patch_equals_full_selective_kernel(acx, torch_gpu, monkeypatch, walls):
    """The selective step (linear pass + CWENO on flagged tiles) equals the
    CWENO kernel run over every tile with the same flags, to FP32 rounding;
    also with walls (a reflective x axis and a transmissive z axis: the
    kernels' wall instantiations)."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    N = 40
    bc = (acx.REFLECTIVE, acx.PERIODIC, acx.TRANSMISSIVE) if walls else (acx.PERIODIC,) * 3
    g = acx.Grid(n=(N, N + 4, N - 4), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85, bc=bc)
    U = torch.from_numpy(learn.steep_pulse_field(g, 21)).cuda()
    rng = np.random.default_rng(4)
    fl = np.zeros(g.ncells, np.uint8)
    idx = rng.choice(g.ncells, g.ncells // 50, replace=False)
    fl[idx] = rng.integers(1, 16, idx.size)
    flags = torch.from_numpy(fl).cuda()
    dt = acx.cfl_dt(g, 0.4)
    a, b = acx.CompensatedState(g, U), acx.CompensatedState(g, U)
    for _ in range(3):
        a.step(dt, flags=flags)
    monkeypatch.setenv("ACX_SEL_FULL", "1")
    for _ in range(3):
        b.step(dt, flags=flags)
    monkeypatch.delenv("ACX_SEL_FULL")
    ra, rb = a.value64().cpu().numpy(), b.value64().cpu().numpy()
    err = np.linalg.norm(ra - rb) / np.linalg.norm(rb)
    print(f"[case] tile patch vs full selective kernel (walls {walls}): {err:.2e}", flush=True)
    assert err < 1e-6


def test_tile_mask_and_masked_step(acx, torch_gpu):
    """acx_sel_mask marks a patch tile (30 x 14 x zc) exactly when it holds a
    flagged cell or a periodic face neighbour of one, and lists the marked
    tiles in ascending order after the mask; the step with that mask is
    bitwise the step that makes the mask itself."""
    import os
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    g = acx.Grid(n=(47, 33, 21), h=1.0 / 47, gamma_hi=0.85, gamma_lo=0.85)
    rng = np.random.default_rng(12)
    fl = np.zeros(g.ncells, np.uint8)
    idx = rng.choice(g.ncells, 40, replace=False)
    fl[idx] = rng.integers(1, 16, idx.size)
    flags = torch.from_numpy(fl).cuda()
    buf = acx.sel_mask(g, flags).cpu().numpy()
    nx, ny, nz = g.n
    ntx, nty = -(-nx // 30), -(-ny // 14)
    zc = min(max(int(os.environ.get("ACX_ZC_PATCH", "8")), 1), nz)   # the tiling's z-chunk
    ntiles = ntx * nty * -(-nz // zc)
    off = (ntiles + 3) // 4 * 4
    assert buf.size == off + 4 * (ntiles + 1)
    mask = buf[:ntiles]
    tl = buf[off:].view(np.int32)
    assert tl[0] == int(mask.sum()) and np.array_equal(tl[1:1 + tl[0]], np.nonzero(mask)[0])
    want = np.zeros(mask.size, np.uint8)
    f3 = fl.reshape(nz, ny, nx)
    for k, j, i in zip(*np.nonzero(f3)):
        for dx, dy, dz in ((0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
            x, y, z = (i + dx) % nx, (j + dy) % ny, (k + dz) % nz
            want[((z // zc) * nty + y // 14) * ntx + x // 30] = 1
    print(f"[case] tile mask {mask.size} tiles, zc {zc}, marked {int(mask.sum())}", flush=True)
    assert np.array_equal(mask, want)
    U = torch.from_numpy(learn.steep_pulse_field(g, 5)).cuda()
    dt = acx.cfl_dt(g, 0.4)
    a, b = acx.CompensatedState(g, U), acx.CompensatedState(g, U)
    mt = torch.from_numpy(buf).cuda()
    for _ in range(2):
        a.step(dt, flags=flags)
        b.step(dt, flags=flags, mask=mt)
    assert torch.equal(a.hi, b.hi) and torch.equal(a.lo, b.lo)


def test_tile_list_over_several_chunks(acx, torch_gpu):
    """The list of marked tiles over a mask of 10 x 20 x 8 = 1600 tiles (the
    listing block walks it in chunks of 1024): the count and the ascending
    indices of the marked entries, for a sparse and a dense marking."""
    import os
    torch = torch_gpu
    g = acx.Grid(n=(300, 280, 64), h=1.0 / 300, gamma_hi=0.85, gamma_lo=0.85)
    zc = min(max(int(os.environ.get("ACX_ZC_PATCH", "8")), 1), 64)
    ntiles = 10 * 20 * -(-64 // zc)
    off = (ntiles + 3) // 4 * 4
    rng = np.random.default_rng(77)
    for frac in (2e-6, 3e-4):
        fl = np.zeros(g.ncells, np.uint8)
        idx = rng.choice(g.ncells, max(1, int(frac * g.ncells)), replace=False)
        fl[idx] = 1
        buf = acx.sel_mask(g, torch.from_numpy(fl).cuda()).cpu().numpy()
        assert buf.size == off + 4 * (ntiles + 1)
        mask = buf[:ntiles]
        tl = buf[off:].view(np.int32)
        print(f"[case] tile list: {ntiles} tiles, {int(mask.sum())} marked", flush=True)
        assert tl[0] == int(mask.sum())
        assert np.array_equal(tl[1:1 + tl[0]], np.nonzero(mask)[0])


def test_patch_list_longer_than_the_grid_of_blocks(acx, torch_gpu, monkeypatch):
    """Every cell flagged on a grid of 9 x 10 x 8 = 720 patch tiles, more than
    the resident blocks of the patch launch (2 per CU), so blocks take several
    list entries each: every cell's output then comes from the patch, and the
    step is bitwise the CWENO kernel run over every tile."""
    import os
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    g = acx.Grid(n=(256, 140, 64), h=1.0 / 256, gamma_hi=0.85, gamma_lo=0.85)
    U = torch.from_numpy(learn.steep_pulse_field(g, 23)).cuda()
    rng = np.random.default_rng(6)
    flags = torch.from_numpy(rng.integers(1, 16, g.ncells).astype(np.uint8)).cuda()
    mask = acx.sel_mask(g, flags)
    ntiles = 9 * 10 * -(-64 // min(max(int(os.environ.get("ACX_ZC_PATCH", "8")), 1), 64))
    off = (ntiles + 3) // 4 * 4
    count = int(mask[off:off + 4].view(torch.int32).item())
    print(f"[case] all cells flagged: {count} of {ntiles} patch tiles listed", flush=True)
    assert count == ntiles and ntiles > 2 * torch.cuda.get_device_properties(0).multi_processor_count
    dt = acx.cfl_dt(g, 0.4)
    a, b = acx.CompensatedState(g, U), acx.CompensatedState(g, U)
    for _ in range(2):
        a.step(dt, flags=flags, mask=mask)
    monkeypatch.setenv("ACX_SEL_FULL", "1")
    for _ in range(2):
        b.step(dt, flags=flags)
    monkeypatch.delenv("ACX_SEL_FULL")
    assert torch.equal(a.hi, b.hi) and torch.equal(a.lo, b.lo)


def test_judge_catches_a_front_that_leaves_its_flags(acx, torch_gpu):
    """Full WENO judges every window of selective steps.  With a network that
    flags nothing (threshold 1) and no dilation, the flags of a window are
    only the cells the weights flag at its start; the steep front moves out
    of them within the 12 steps, the judge finds cells the window ran with
    the linear polynomial whose CWENO4 face averages differ from it, and
    reruns the window with full CWENO4.  The main run
    therefore equals the all-WENO FP32 stepper bit for bit, while the same
    run without the judge does not."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    N = 48
    g = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    U0 = torch.from_numpy(learn.steep_pulse_field(g, 77)).cuda()
    dt = acx.cfl_dt(g, 0.4)
    packed = learn.pack(torch.from_numpy(learn.init_params(0)).cuda())
    judged = learn.SelectiveStepper(g, packed, U0, thr=1.0, refresh=12, dilate=0)
    blind = learn.SelectiveStepper(g, packed, U0, thr=1.0, refresh=12, dilate=0, judge=False)
    full = acx.CompensatedState(g, U0)
    for _ in range(24):
        judged.step(dt)
        blind.step(dt)
        full.step(dt)
    misses = judged.finish()
    print(f"[case] judge: {misses} misses (largest skipped face difference {judged.worst_skipped:.2e}, "
          f"tolerance {judged.judge_tol:.2e}), {judged.redone_windows} of {judged.judged_windows} windows rerun",
          flush=True)
    assert misses > 0 and judged.redone_windows == 2
    assert torch.equal(judged.state.hi, full.hi) and torch.equal(judged.state.lo, full.lo)
    e_blind = acx.diff_f32c(blind.state.hi, blind.state.lo, full.value64())[0]
    print(f"[case] without the judge: rel L2 {e_blind:.2e} against all-WENO", flush=True)
    assert e_blind > 1e-7


def test_judge_labels_match_the_fp64_labels(acx, torch_gpu):
    """The judge's FP32 labels agree with the FP64 label kernel (the training
    labels) except where the weight deviation sits within FP32 rounding of
    the threshold."""
    torch = torch_gpu
    from acoustic_petsc_amd import learn
    N = 40
    g = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=0.85, gamma_lo=0.85)
    U = torch.from_numpy(learn.steep_pulse_field(g, 8)).cuda()
    lab64, dev = learn.labels(g, U, 0.05)
    out, _, _ = learn.judge(g, U.to(torch.float32))
    bits = torch.stack([(out >> v) & 1 for v in range(4)], dim=1).reshape(-1)
    differ = bits != lab64
    near = (dev - 0.05).abs() < 1e-4
    print(f"[case] judge vs FP64 labels: {int(differ.sum())} differ, {int(lab64.sum())} labelled", flush=True)
    assert int(lab64.sum()) > 0
    assert bool((~differ | near).all())
