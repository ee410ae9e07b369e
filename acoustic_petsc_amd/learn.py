"""Learned troubled-cell indicator: data from the solver's own weights,
training with the library's MFMA forward / fp32 backward kernels, and the
selective FP32 stepper that uses it (include/acx_learn.h).

No counterpart in the reference (it always evaluates the full CWENO4 weights,
3d/weno4/plane_wave/reconstruction.c:38-259).  The label of a (cell,
component) is sum_k |w_k - gamma_k| > 0.05 over the eight normalised CWENO
weights (R4 and seven R3 stencils, reconstruction.c:205-218) with the
nonlinear linear-weight setting gamma = (0.85, 0.85)
(Viscous-Multiphase-Fourth-Order/reconstruction.c:42-43).  With w = gamma the
CWENO combination (reconstruction.c:223-257) is exactly the R4 polynomial.
"""
import ctypes as _C
import math

import numpy as np

from . import _check, _ptr, _stream, lib, ACOUSTIC4, Grid, CompensatedState, field_noise, sel_mask

NPARAM = 32 * 32 + 32 + 32 * 32 + 32 + 32 + 1
NFEAT = 19
OFF = {"w1": 0, "b1": 1024, "w2": 1056, "b2": 2080, "w3": 2112, "b3": 2144}
# hidden activation codes of acx_mlp_pack_act / acx_mlp_backward_act
ACT = {"relu": 0, "tanh": 1}


def _bind():
    L = lib()
    if getattr(L, "_learn_bound", False):
        return L
    vp, i32, f32, dbl = _C.c_void_p, _C.c_int, _C.c_float, _C.c_double
    gp = _C.c_void_p
    for name, args in {
        "acx_mlp_pack": [vp, vp, vp],
        "acx_mlp_pack_act": [vp, i32, vp, vp],
        "acx_mlp_backward_act": [vp, i32, vp, vp, vp, vp, vp, i32, f32, vp, vp],
        "acx_features_f32": [gp, vp, vp, vp],
        "acx_weno_labels_f64": [gp, vp, dbl, vp, vp, vp],
        "acx_mlp_forward": [vp, vp, i32, vp, vp, vp, vp],
        "acx_mlp_backward": [vp, vp, vp, vp, vp, vp, i32, f32, vp, vp],
        "acx_mlp_adam": [vp, vp, vp, vp, i32, f32, f32, f32, f32, vp],
        "acx_indicator_flags": [gp, vp, vp, vp, f32, vp, vp],
        "acx_flags_dilate": [gp, vp, vp, vp],
        "acx_flags_count": [vp, _C.c_size_t, _C.POINTER(_C.c_longlong), vp],
        "acx_weno_judge_f32": [gp, vp, vp, vp, dbl, dbl, vp, vp, vp, vp],
    }.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L.acx_mlp_packed_bytes.restype = _C.c_size_t
    L._learn_bound = True
    return L


def _g(grid):
    g = grid.c_struct()
    return g, _C.cast(_C.pointer(g), _C.c_void_p)


def _act(act):
    if act not in ACT:
        raise ValueError(f"activation must be one of {sorted(ACT)}, not {act!r}")
    return ACT[act]


def init_params(seed=0, act="relu"):
    """Zero biases and uniform weights (fp32 numpy, NPARAM): He limits
    sqrt(6 / fan_in) for ReLU layers, Glorot sqrt(6 / (fan_in + fan_out))
    for tanh layers."""
    _act(act)
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    if act == "tanh":
        lim1, lim2, lim3 = math.sqrt(6.0 / (NFEAT + 32)), math.sqrt(6.0 / 64), math.sqrt(6.0 / 33)
    else:
        lim1, lim2 = math.sqrt(6.0 / NFEAT), math.sqrt(6.0 / 32)
        lim3 = lim2
    w1 = rng.uniform(-lim1, lim1, (32, 32)).astype(np.float32)
    w1[:, NFEAT:] = 0.0
    p[OFF["w1"]:OFF["w1"] + 1024] = w1.ravel()
    p[OFF["w2"]:OFF["w2"] + 1024] = rng.uniform(-lim2, lim2, 1024)
    p[OFF["w3"]:OFF["w3"] + 32] = rng.uniform(-lim3, lim3, 32)
    return p


def pack(params_dev, stream=None, act="relu"):
    """MFMA operand layout of the parameters; it carries the hidden
    activation, which forward, the fused indicator and the stepper use."""
    import torch
    L = _bind()
    buf = torch.empty(L.acx_mlp_packed_bytes(), dtype=torch.uint8, device="cuda")
    _check(L.acx_mlp_pack_act(_ptr(params_dev, torch.float32), _act(act), _C.c_void_p(buf.data_ptr()),
                              _stream(stream)))
    return buf


def features(grid, U32, stream=None):
    import torch
    L = _bind()
    X = torch.empty((grid.ncells * 4, 32), dtype=torch.float32, device="cuda")
    g, gp = _g(grid)
    _check(L.acx_features_f32(gp, _ptr(U32, torch.float32), _ptr(X, torch.float32), _stream(stream)))
    return X


def labels(grid, U64, thr=0.05, stream=None):
    """(labels uint8 [ncells*4], weight deviation float32 [ncells*4]) with the grid's gammas."""
    import torch
    L = _bind()
    T = torch.empty(grid.ncells * 4, dtype=torch.uint8, device="cuda")
    W = torch.empty(grid.ncells * 4, dtype=torch.float32, device="cuda")
    g, gp = _g(grid)
    _check(L.acx_weno_labels_f64(gp, _ptr(U64, torch.float64), thr, _C.c_void_p(T.data_ptr()),
                                 _ptr(W, torch.float32), _stream(stream)))
    return T, W


def forward(packed, X, keep=False, stream=None):
    import torch
    L = _bind()
    n = X.shape[0]
    Y = torch.empty(n, dtype=torch.float32, device="cuda")
    H1 = torch.empty((n, 32), dtype=torch.float32, device="cuda") if keep else None
    H2 = torch.empty((n, 32), dtype=torch.float32, device="cuda") if keep else None
    _check(L.acx_mlp_forward(_C.c_void_p(packed.data_ptr()), _ptr(X, torch.float32), n, _ptr(Y, torch.float32),
                             _ptr(H1, torch.float32) if keep else None, _ptr(H2, torch.float32) if keep else None,
                             _stream(stream)))
    return (Y, H1, H2) if keep else Y


def backward(params_dev, X, H1, H2, Y, T, pos_weight, stream=None, act="relu"):
    import torch
    L = _bind()
    G = torch.empty(NPARAM, dtype=torch.float32, device="cuda")
    _check(L.acx_mlp_backward_act(_ptr(params_dev, torch.float32), _act(act), _ptr(X, torch.float32),
                                  _ptr(H1, torch.float32), _ptr(H2, torch.float32), _ptr(Y, torch.float32),
                                  _C.c_void_p(T.data_ptr()), X.shape[0], pos_weight, _ptr(G, torch.float32),
                                  _stream(stream)))
    return G


def steep_pulse_field(grid, seed, noise_rms=1e-4):
    """Spherical steep-fronted pressure pulse p = A (1 - tanh((r - r0)/(1.5 h)))/2,
    A ~ U[1e-3, 1e-2], r0 ~ U[0.05, 0.15], centre U[0,1)^3, over smooth
    band-limited noise (|k| in [1, 4]); zero velocity; point values at cell
    centres (numpy, (nz, ny, nx, 4))."""
    rng = np.random.default_rng(seed)
    nz, ny, nx, _ = grid.shape()
    A = rng.uniform(1e-3, 1e-2)
    r0 = rng.uniform(0.05, 0.15)
    c = rng.uniform(0, 1, 3)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz

    def mi(a, c0):
        d = a - c0
        return d - np.round(d)
    dx2 = mi(x, c[0])[None, None, :] ** 2
    dy2 = mi(y, c[1])[None, :, None] ** 2
    dz2 = mi(z, c[2])[:, None, None] ** 2
    r = np.sqrt(dx2 + dy2 + dz2)
    h = 1.0 / nx
    U = field_noise(grid, seed + 7, 1.0, 4.0, noise_rms, 0.0, 0.05)
    U[..., 0] += A * 0.5 * (1.0 - np.tanh((r - r0) / (1.5 * h)))
    return U


def make_dataset(n_snapshots=8, N=64, seed0=100, steps_between=12, thr=0.05):
    """Solver-labelled samples from N^3 runs (one run per snapshot seed): the
    FP64 nonlinear scheme advances a steep pulse; every `steps_between` steps
    the state is labelled with its own CWENO weights.  Returns (X, T) on the
    GPU: X float32 [n, 32], T uint8 [n]."""
    import torch
    from . import ssp_rks3_step, cfl_dt
    grid = Grid(n=(N, N, N), h=1.0 / N, system=ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    Xs, Ts = [], []
    per_run = 2
    runs = (n_snapshots + per_run - 1) // per_run
    dt = cfl_dt(grid, 0.4)
    for r in range(runs):
        U = torch.from_numpy(steep_pulse_field(grid, seed0 + r)).cuda()
        work = (torch.empty_like(U), torch.empty_like(U))
        for k in range(per_run):
            if len(Xs) >= n_snapshots:
                break
            for _ in range(steps_between if k else 1):
                ssp_rks3_step(grid, U, dt, work)
            T, _ = labels(grid, U, thr)
            Xs.append(features(grid, U.to(torch.float32)))
            Ts.append(T)
    return torch.cat(Xs), torch.cat(Ts)


def train(X, T, epochs=8, batch=8192, lr=3e-3, seed=0, pos_weight=None, log=None, act="relu"):
    """Trains the indicator with the library's kernels; returns fp32 params
    (GPU), to be packed with the same activation."""
    import torch
    L = _bind()
    n = X.shape[0]
    p = torch.from_numpy(init_params(seed, act)).cuda()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    if pos_weight is None:
        npos = max(int(T.sum().item()), 1)
        pos_weight = min(50.0, (n - npos) / npos)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=gen)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            xb, tb = X[idx].contiguous(), T[idx].contiguous()
            packed = pack(p, act=act)
            Y, H1, H2 = forward(packed, xb, keep=True)
            G = backward(p, xb, H1, H2, Y, tb, pos_weight, act=act)
            step += 1
            _check(L.acx_mlp_adam(_ptr(p, torch.float32), _ptr(G, torch.float32), _ptr(m, torch.float32),
                                  _ptr(v, torch.float32), step, lr, 0.9, 0.999, 1e-8, _stream()))
            if log is not None:
                with torch.no_grad():
                    y = Y.clamp(1e-7, 1 - 1e-7)
                    t = tb.float()
                    c = torch.where(tb > 0, torch.full_like(t, pos_weight), torch.ones_like(t))
                    tot += float((-(c * (t * torch.log(y) + (1 - t) * torch.log(1 - y)))).sum())
        if log is not None:
            log.append(tot / n)
    return p


def evaluate(packed, X, T, thr):
    """(recall, precision, flagged fraction) of y > thr against labels T."""
    Y = forward(packed, X)
    pred = Y > thr
    t = T > 0
    tp = int((pred & t).sum())
    return tp / max(int(t.sum()), 1), tp / max(int(pred.sum()), 1), float(pred.float().mean())


def judge(grid, hi, used=None, base=None, out=None, thr=0.05, tol=0.0, stream=None):
    """Full-WENO judge (acx_weno_judge_f32) of the FP32 state `hi`: returns
    (out = base | labels, misses, worst).  With `used` (the flags a run of
    selective steps ran with), misses counts the (cell, component) pairs run
    with the linear polynomial whose CWENO4 face averages differ from the
    linear ones by more than `tol`, and worst is the largest such difference;
    without it both are None."""
    import torch
    L = _bind()
    g, gp = _g(grid)
    if out is None:
        out = torch.empty(grid.ncells, dtype=torch.uint8, device="cuda")
    missed, worst = _C.c_longlong(0), _C.c_double(0.0)
    vp = lambda t: _C.c_void_p(t.data_ptr()) if t is not None else None
    judged = used is not None
    _check(L.acx_weno_judge_f32(gp, _ptr(hi, torch.float32), vp(used), vp(base), thr, tol, vp(out),
                                _C.byref(missed) if judged else None, _C.byref(worst) if judged else None,
                                _stream(stream)))
    return out, (missed.value if judged else None), (worst.value if judged else None)


class SelectiveStepper:
    """FP32 compensated SSP-RK3 with the learned indicator: flags from the
    MLP every `refresh` steps, dilated by `dilate` cells and spread over the
    four components (a cell flagged for any component is flagged for all),
    full CWENO4 where set.

    Full WENO judges every window of `refresh` steps: at the next refresh the
    full CWENO4 reconstruction of the state reached is compared, at every
    (cell, component) the window ran with the linear polynomial, with the
    linear one (acx_weno_judge_f32: largest face-average difference).  A
    difference above judge_rel x max|U| (max over the initial state) is a
    miss: the window is run again from its saved start state with full CWENO4
    everywhere (the fallback).  The default 1e-6 sits above the FP32 rounding
    of the face values the judge compares (about 1e-7 of max|U|); a face
    difference d left in a cell for a 12-step window at CFL 0.4 moves its
    state by about 12 x 0.4/3 x d = 1.6 d, so an accepted window changes no
    cell by more than ~2e-6 of max|U|.  The counts are kept in `misses`,
    `redone_windows` and `judged_windows`, the largest difference seen in
    `worst_skipped`.  The new flags are the MLP's united with the cells whose
    weights leave the linear weights now (the training label), before the
    dilation.  judge=False switches the check off.

    Defaults: at CFL 0.4 a front moves 0.4/3 cell per step, so flags taken
    every 12 steps and dilated by 3 cells still cover it (1.6 cells) with a
    margin of more than a cell; the indicator and the judge then run once per
    12 steps."""

    def __init__(self, grid, packed, U64, thr=0.3, refresh=12, dilate=3, judge=True, label_thr=0.05,
                 judge_rel=1e-6):
        import torch
        self.grid, self.packed, self.thr, self.refresh, self.dilate = grid, packed, thr, refresh, dilate
        self.use_judge, self.label_thr = judge, label_thr
        self.judge_tol = judge_rel * float(U64.abs().max())
        self.worst_skipped = 0.0
        self.state = CompensatedState(grid, U64)
        self.flags = torch.empty(grid.ncells, dtype=torch.uint8, device="cuda")
        self.tmp = torch.empty_like(self.flags)
        self.tmp2 = torch.empty_like(self.flags)
        self.mask = None
        self.nsteps = 0
        self.flag_bits = []
        self.misses = 0
        self.redone_windows = 0
        self.judged_windows = 0
        self._saved = None       # (hi, lo) at the start of the current window
        self._swap_in = None     # the free (hi, lo) pair that replaces the saved one after a window's first step
        self._window_dts = []

    def _mlp_flags(self, stream=None):
        L = _bind()
        g, gp = _g(self.grid)
        import torch
        _check(L.acx_indicator_flags(gp, _ptr(self.state.hi, torch.float32), None,
                                     _C.c_void_p(self.packed.data_ptr()), self.thr,
                                     _C.c_void_p(self.tmp.data_ptr()), _stream(stream)))

    def _dilate_into_flags(self, src, stream=None):
        L = _bind()
        g, gp = _g(self.grid)
        spare = self.tmp if src is not self.tmp else self.tmp2
        dst = self.flags
        for _ in range(self.dilate):
            _check(L.acx_flags_dilate(gp, _C.c_void_p(src.data_ptr()), _C.c_void_p(dst.data_ptr()), _stream(stream)))
            src, dst = dst, (spare if dst is self.flags else self.flags)
        if src is not self.flags:
            self.flags.copy_(src)

    def refresh_flags(self, stream=None):
        """Judge the finished window (rerun it with full WENO on a miss), then
        take new flags from the MLP and the current weights."""
        import torch
        have_window = self._saved is not None and self._window_dts
        if self.use_judge and have_window:
            self._mlp_flags(stream)
            _, missed, worst = judge(self.grid, self.state.hi, used=self.flags, base=self.tmp, out=self.tmp2,
                                     thr=self.label_thr, tol=self.judge_tol, stream=stream)
            self.judged_windows += 1
            self.worst_skipped = max(self.worst_skipped, worst)
            if missed:
                self.misses += missed
                self.redone_windows += 1
                self.state.hi.copy_(self._saved[0])
                self.state.lo.copy_(self._saved[1])
                for dt in self._window_dts:
                    self.state.step(dt, stream=stream)      # full CWENO4 everywhere
                self._mlp_flags(stream)
                judge(self.grid, self.state.hi, base=self.tmp, out=self.tmp2, thr=self.label_thr, stream=stream)
            src = self.tmp2
        else:
            self._mlp_flags(stream)
            if self.use_judge:
                judge(self.grid, self.state.hi, base=self.tmp, out=self.tmp2, thr=self.label_thr, stream=stream)
                src = self.tmp2
            else:
                src = self.tmp
        self._dilate_into_flags(src, stream)
        # the patch tiles of these flags, kept until the next refresh
        self.mask = sel_mask(self.grid, self.flags, self.mask, stream)
        if self.use_judge:
            # The window's start state is kept without a copy: the pair that
            # holds it now stays out of the state's ping-pong until the next
            # refresh (the first step reads it, and its place as the next
            # output is taken by the free pair); three pairs in all.
            old = self._saved
            self._saved = (self.state.hi, self.state.lo)
            self._swap_in = old if old is not None else (torch.empty_like(self.state.hi),
                                                        torch.empty_like(self.state.lo))
        self._window_dts = []

    def step(self, dt):
        if self.nsteps % self.refresh == 0:
            self.refresh_flags()
        self.state.step(dt, flags=self.flags, mask=self.mask)
        if self._swap_in is not None:
            # the step wrote the free pair's partner; the saved pair would be the next output
            self.state._hi2, self.state._lo2 = self._swap_in
            self._swap_in = None
        self._window_dts.append(dt)
        self.nsteps += 1

    def finish(self):
        """Judge the last (partial) window; returns the total misses."""
        if self.use_judge and self._window_dts:
            self.refresh_flags()
        return self.misses

    def flagged_fraction(self):
        L = _bind()
        bits = _C.c_longlong()
        _check(L.acx_flags_count(_C.c_void_p(self.flags.data_ptr()), self.grid.ncells, _C.byref(bits), _stream()))
        return bits.value / (4.0 * self.grid.ncells)
