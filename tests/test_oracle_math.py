"""Closed-form and invariant checks that pin the oracle's restatement.

Each check is derived from the mathematics, not from the oracle's formulas:
polynomial reproduction (exact cell averages by Gauss quadrature), the
smoothness indicator as a sum of squared derivative integrals (sympy),
linear weights recovered on quadratic data, reflection symmetry, the
stability polynomial and order of the SSP steppers, telescoping
conservation, energy dissipation of the LLF flux, and 4th-order convergence
of the acoustic operator to an exact plane wave.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest
import sympy as sp

X, Y, Z = sp.symbols("x y z")
R = sp.Rational
BASIS3 = [1, X, Y, Z, X * X - R(1, 12), Y * Y - R(1, 12), Z * Z - R(1, 12), X * Y, Y * Z, Z * X,
          X * (X * X - R(3, 20)), Y * (Y * Y - R(3, 20)), Z * (Z * Z - R(3, 20)),
          (X * X - R(1, 12)) * Y, X * (Y * Y - R(1, 12)), (Y * Y - R(1, 12)) * Z, Y * (Z * Z - R(1, 12)),
          (X * X - R(1, 12)) * Z, X * (Z * Z - R(1, 12)), X * Y * Z]
HALF = R(1, 2)


def cell_averages_3d(p):
    """Exact cell averages of a polynomial (degree <= 5 per variable) over the
    5x5x5 block of unit cells centred on the origin, indexed [z][y][x]."""
    f = sp.lambdify((X, Y, Z), p, "numpy")
    gx, gw = np.polynomial.legendre.leggauss(3)
    gx, gw = gx / 2, gw / 2
    W = gw[:, None, None] * gw[None, :, None] * gw[None, None, :]
    blk = np.zeros((5, 5, 5))
    for k in range(5):
        for j in range(5):
            for i in range(5):
                A, B, Cc = np.meshgrid(i - 2 + gx, j - 2 + gx, k - 2 + gx, indexing="ij")
                val = f(A, B, Cc)
                blk[k, j, i] = np.sum(W * val) if np.ndim(val) else float(val)
    return blk


def derivative_energy(p, max_order):
    """sum over multi-indices 1 <= |a| <= max_order of the cell integral of (D^a p)^2."""
    tot = 0
    for order in range(1, max_order + 1):
        for combo in itertools.combinations_with_replacement((X, Y, Z), order):
            d = sp.diff(p, *combo)
            tot += sp.integrate(d ** 2, (X, -HALF, HALF), (Y, -HALF, HALF), (Z, -HALF, HALF))
    return float(tot)


def random_poly(rng, degree):
    n = {1: 4, 2: 10, 3: 20}[degree]
    c = np.zeros(20)
    c[:n] = rng.standard_normal(n)
    return c, sum(float(ci) * b for ci, b in zip(c, BASIS3))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cweno3d_linear_reproduces_cubics(orc, seed):
    c, p = random_poly(np.random.default_rng(seed), 3)
    co, dg = orc.cweno3d(cell_averages_3d(p), (1.0, 0.0))
    assert np.max(np.abs(co - c)) < 1e-12
    assert dg[0] == 1.0   # wt_ratio of the shipped weights is exactly one


def test_is_r4_is_derivative_energy(orc):
    c, p = random_poly(np.random.default_rng(11), 3)
    _, dg = orc.cweno3d(cell_averages_3d(p), (0.85, 0.85))
    assert abs(dg[9] - derivative_energy(p, 3)) <= 1e-12 * abs(dg[9])


@pytest.mark.parametrize("gamma", [(0.85, 0.85), (0.5, 0.3), (0.9, 0.1)])
def test_quadratic_data_gives_linear_weights(orc, gamma):
    """On quadratic data every sub-stencil is exact, all indicators agree and
    the nonlinear weights equal the linear ones; the output is the exact
    polynomial."""
    c, p = random_poly(np.random.default_rng(5), 2)
    co, dg = orc.cweno3d(cell_averages_3d(p), gamma)
    g4 = gamma[0]
    g3 = [(1 - gamma[0]) * gamma[1]] + [(1 - gamma[0]) * (1 - gamma[1]) / 6] * 6
    assert abs(dg[1] - g4) < 1e-12
    assert np.max(np.abs(np.array(dg[2:9]) - g3)) < 1e-12
    assert abs(sum(dg[1:9]) - 1.0) < 1e-14
    assert np.max(np.abs(co - c)) < 1e-12


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("scale", [1.0, 1e-6])
def test_sided_weights_follow_the_indicator_ratio(orc, scale, axis):
    """A block that is one quadratic q1 on the a >= 0 side of an axis a and
    another, q2 = q1 + b a + c (a^2 - 1/12), on the a <= 0 side (the two have
    the same cell averages on the a = 0 plane).  The +a stencil reads only
    a >= 0 cells and the -a stencil only a <= 0 cells, so each reproduces its
    quadratic and its indicator is that quadratic's derivative energy (sympy,
    orders 1 and 2).  With the published weights gamma / (IS + eps)^4,
    eps = 1e-12 (reconstruction.c:205-213, hype.h:37), and equal gammas of the
    two sided stencils, w(+a) / w(-a) = ((IS(q2) + eps) / (IS(q1) + eps))^4.
    At unit scale this pins the exponent and the indicator's coefficients; at
    1e-6 (IS near eps) it pins eps."""
    A, ax, wp = {"x": (X, 2, 3), "y": (Y, 1, 5), "z": (Z, 0, 7)}[axis]   # symbol, block axis, diag of w(+a)
    rng = np.random.default_rng(21)
    _, q1 = random_poly(rng, 2)
    b, c = rng.standard_normal(2)
    q1 = scale * q1
    q2 = q1 + scale * (b * A + c * (A * A - R(1, 12)))
    blk = cell_averages_3d(q2)
    side = [slice(None)] * 3
    side[ax] = slice(2, None)
    blk[tuple(side)] = cell_averages_3d(q1)[tuple(side)]
    _, dg = orc.cweno3d(blk, (0.85, 0.85))
    e1, e2 = derivative_energy(q1, 2), derivative_energy(q2, 2)
    want = ((e2 + 1e-12) / (e1 + 1e-12)) ** 4
    got = dg[wp] / dg[wp + 1]   # w(+a) / w(-a)
    assert abs(got / want - 1.0) < 1e-9, (got, want, e1, e2)
    # and the sides are told apart: the steeper side gets the smaller weight
    assert (got > 1.0) == (e2 > e1)


def test_weights_avoid_a_discontinuity(orc):
    """A jump between the centre cell and its +x neighbour: the -x sided
    stencil must dominate the +x sided one, and the 4th-order weight collapses."""
    blk = np.zeros((5, 5, 5))
    blk[:, :, 3:] = 1.0
    _, dg = orc.cweno3d(blk, (0.85, 0.85))
    w_px, w_mx = dg[3], dg[4]
    assert w_mx > 100 * w_px
    assert dg[1] < 1e-6
    assert abs(sum(dg[1:9]) - 1.0) < 1e-14


def test_mirror_symmetry(orc):
    rng = np.random.default_rng(9)
    blk = rng.standard_normal((5, 5, 5))
    co, dg = orc.cweno3d(blk, (0.85, 0.85))
    cm, dm = orc.cweno3d(blk[:, :, ::-1].copy(), (0.85, 0.85))
    odd_x = [1, 7, 9, 10, 14, 18, 19]
    sign = np.array([-1.0 if k in odd_x else 1.0 for k in range(20)])
    assert np.max(np.abs(cm - sign * co)) < 1e-13 * np.max(np.abs(co))
    assert abs(dm[3] - dg[4]) < 1e-13 and abs(dm[4] - dg[3]) < 1e-13


def test_cweno2d_reproduces_quadratics_and_cubics(orc):
    rng = np.random.default_rng(3)
    b2 = [1, X, Y, X * X - R(1, 12), Y * Y - R(1, 12), X * Y, X * (X * X - R(3, 20)), Y * (Y * Y - R(3, 20)),
          Y * (X * X - R(1, 12)), X * (Y * Y - R(1, 12))]
    for deg, n in ((2, 6), (3, 10)):
        c = np.zeros(10)
        c[:n] = rng.standard_normal(n)
        p = sum(float(ci) * b for ci, b in zip(c, b2))
        f = sp.lambdify((X, Y), p, "numpy")
        gx, gw = np.polynomial.legendre.leggauss(3)
        gx, gw = gx / 2, gw / 2
        blk = np.zeros((5, 5))
        for j in range(5):
            for i in range(5):
                A, B = np.meshgrid(i - 2 + gx, j - 2 + gx, indexing="ij")
                v = f(A, B)
                blk[j, i] = np.sum(gw[:, None] * gw[None, :] * v) if np.ndim(v) else float(v)
        lin = np.zeros(10)
        orc.lib().orc_weno2d_linear(C.cast(blk.ctypes.data + 8 * 12, orc.dp), 1, 5, orc.P(lin))
        assert np.max(np.abs(lin - c)) < 1e-12
        if deg == 2:
            nl = np.zeros(10)
            orc.lib().orc_cweno2d(C.cast(blk.ctypes.data + 8 * 12, orc.dp), 1, 5, 0.85, 0.85, orc.P(nl))
            assert np.max(np.abs(nl - c)) < 1e-12


# ---- steppers ---------------------------------------------------------------

RHS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))


def _linear_ode(a, b):
    """u' = A u with A = [[a, -b], [b, a]] (eigenvalues a +- ib)."""
    def f(ctx, t, U, F):
        F[0] = a * U[0] - b * U[1]
        F[1] = b * U[0] + a * U[1]
        return 0
    return RHS_FN(f)


def _step_map(orc, fn, scheme, dt):
    cols = []
    for e in ([1.0, 0.0], [0.0, 1.0]):
        u = np.array(e)
        w0, w1, F = np.zeros(2), np.zeros(2), np.zeros(2)
        if scheme == 0:
            orc.lib().orc_ssp_rks3_step(C.cast(fn, C.c_void_p), None, 2, orc.P(u), 0.0, dt, 4,
                                        orc.P(w0), orc.P(w1), orc.P(F))
        else:
            orc.lib().orc_ssp_rk104_step(C.cast(fn, C.c_void_p), None, 2, orc.P(u), 0.0, dt,
                                         orc.P(w0), orc.P(w1), orc.P(F))
        cols.append(u)
    return np.array(cols).T


def test_rks3_stability_polynomial(orc):
    """One TSSSPRKS3 (s = 4) step of u' = lambda u multiplies by
    R(z) = 1 + z + z^2/2 + z^3/6 + z^4/48."""
    for a, b, dt in ((-0.3, 1.1, 0.7), (0.2, -2.0, 0.5), (-1.5, 0.0, 0.9)):
        z = complex(a, b) * dt
        Rz = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 48
        M = _step_map(orc, _linear_ode(a, b), 0, dt)
        assert abs(M[0, 0] - Rz.real) < 1e-14 and abs(M[1, 0] - Rz.imag) < 1e-14


def test_rk104_is_fourth_order_on_linear_and_nonlinear(orc):
    z = complex(-0.2, 0.9) * 0.1
    M = _step_map(orc, _linear_ode(-0.2, 0.9), 1, 0.1)
    ez = np.exp(z)
    assert abs(complex(M[0, 0], M[1, 0]) - ez) < 2 * abs(z) ** 5
    # nonlinear: u' = -u^2, u(0) = 1, u(t) = 1/(1+t)

    def f(ctx, t, U, F):
        F[0] = -U[0] * U[0]
        F[1] = 0.0
        return 0
    fn = RHS_FN(f)
    errs = []
    for n in (10, 20, 40):
        u = np.array([1.0, 0.0])
        orc.lib().orc_ts_solve(C.cast(fn, C.c_void_p), None, 2, orc.P(u), 0.0, 2.0 / n, 2.0, 1, 0)
        errs.append(abs(u[0] - 1.0 / 3.0))
    assert math.log2(errs[1] / errs[2]) > 3.8


def test_rks3_is_third_order_nonlinear(orc):
    def f(ctx, t, U, F):
        F[0] = -U[0] * U[0]
        F[1] = math.cos(t)   # non-autonomous: checks the stage times 0, 1/2, 1, 1/2
        return 0
    fn = RHS_FN(f)
    errs = []
    for n in (20, 40, 80):
        u = np.array([1.0, 0.0])
        orc.lib().orc_ts_solve(C.cast(fn, C.c_void_p), None, 2, orc.P(u), 0.0, 2.0 / n, 2.0, 0, 4)
        errs.append(abs(u[0] - 1.0 / 3.0) + abs(u[1] - math.sin(2.0)))
    r = math.log2(errs[1] / errs[2])
    assert 2.8 < r < 3.4


def test_matchstep_lands_on_final_time(orc):
    def f(ctx, t, U, F):
        F[0] = 1.0
        F[1] = 0.0
        return 0
    fn = RHS_FN(f)
    u = np.array([0.0, 0.0])
    steps = orc.lib().orc_ts_solve(C.cast(fn, C.c_void_p), None, 2, orc.P(u), 0.0, 0.3, 1.0, 0, 4)
    assert steps == 4 and abs(u[0] - 1.0) < 1e-15


# ---- right-hand sides -------------------------------------------------------

@pytest.mark.parametrize("system,gamma", [(1, (1.0, 0.0)), (1, (0.85, 0.85)), (0, (1.0, 0.0)), (0, (0.85, 0.85))])
def test_rhs_conserves_on_periodic_grid(orc, system, gamma):
    n = (12, 10, 8)
    g = orc.grid3d(n, 0.1, system, gamma)
    U = np.random.default_rng(4).standard_normal((8, 10, 12, orc.nvar(system)))
    F = orc.rhs3d(g, U)
    comps = range(4) if system == 1 else [1]
    for c in comps:
        assert abs(F[..., c].sum()) < 1e-11 * np.abs(F[..., c]).sum()


@pytest.mark.parametrize("system,gamma,bc", [(1, (1.0, 0.0), (0, 0, 0)), (1, (0.85, 0.85), (0, 0, 0)),
                                             (0, (1.0, 0.0), (0, 0, 0)), (0, (0.85, 0.85), (1, 1, 1)),
                                             (0, (1.0, 0.0), (0, 1, 1))])
def test_streaming_equals_faithful(orc, system, gamma, bc):
    n = (13, 11, 9)
    g = orc.grid3d(n, 1.0 / 13, system, gamma, bc=bc)
    U = np.random.default_rng(8).standard_normal((9, 11, 13, orc.nvar(system)))
    a = orc.rhs3d(g, U, faithful=True)
    b = orc.rhs3d(g, U, faithful=False)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


def test_llf_dissipates_acoustic_energy(orc):
    """dE/dt = sum(p F_p / K + rho0 u.F_u) <= 0 for the LLF scheme (the
    central part is skew-adjoint, the upwind part negative semi-definite)."""
    n = (10, 10, 10)
    for gamma in ((1.0, 0.0), (0.85, 0.85)):
        g = orc.grid3d(n, 0.1, orc.ACOUSTIC4, gamma)
        for seed in range(3):
            U = np.random.default_rng(seed).standard_normal((10, 10, 10, 4))
            F = orc.rhs3d(g, U)
            dE = np.sum(U[..., 0] * F[..., 0]) + np.sum(U[..., 1:] * F[..., 1:])
            assert dE < 0


def test_central_wave_operator_conserves_energy_linear(orc):
    """The WAVE2 central-flux operator with the linear reconstruction is
    self-adjoint: sum(u * Lap_h w) = sum(w * Lap_h u)."""
    n = (10, 10, 10)
    g = orc.grid3d(n, 0.1, orc.WAVE2)
    rng = np.random.default_rng(1)
    U, W = rng.standard_normal((10, 10, 10, 2)), rng.standard_normal((10, 10, 10, 2))
    U[..., 1] = 0
    W[..., 1] = 0
    LU, LW = orc.rhs3d(g, U)[..., 1], orc.rhs3d(g, W)[..., 1]
    a, b = np.sum(W[..., 0] * LU), np.sum(U[..., 0] * LW)
    assert abs(a - b) < 1e-12 * abs(a)


def _acoustic_plane_wave(N, k, A, t):
    """Exact cell averages of p = A sin(th) cos(wt), u = -(A/rho c) khat cos(th) sin(wt),
    th = 2 pi k.x (rho0 = c0 = 1, unit periodic cube)."""
    kk = math.sqrt(sum(v * v for v in k))
    w = 2 * math.pi * kk
    damp = np.prod([np.sinc(kd / N) for kd in k])
    i = (np.arange(N) + 0.5) / N
    Zc, Yc, Xc = np.meshgrid(i, i, i, indexing="ij")
    th = 2 * math.pi * (k[0] * Xc + k[1] * Yc + k[2] * Zc)
    U = np.zeros((N, N, N, 4))
    U[..., 0] = A * np.sin(th) * math.cos(w * t) * damp
    for d in range(3):
        U[..., 1 + d] = -A * (k[d] / kk) * np.cos(th) * math.sin(w * t) * damp
    return U


@pytest.mark.slow
def test_acoustic_operator_converges_at_fourth_order(orc):
    """LLF + CWENO4 (shipped weights) on p = sin(2 pi (x+2y+z)): the error at
    T = 0.05 with a small CFL drops by ~2^4 from N = 16 to N = 32."""
    k, A, T = (1, 2, 1), 1e-3, 0.05
    errs = []
    for N in (16, 32):
        g = orc.grid3d(N, 1.0 / N, orc.ACOUSTIC4, (1.0, 0.0))
        U = _acoustic_plane_wave(N, k, A, 0.0)
        dt = 0.05 / N
        orc.ts_solve("orc_rhs3d_cb_stream", g, U, dt, T, 0, 4)
        errs.append(np.linalg.norm(U - _acoustic_plane_wave(N, k, A, T)) / np.linalg.norm(U))
    rate = math.log2(errs[0] / errs[1])
    assert 3.6 < rate < 4.6, (errs, rate)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_reflective_wall_is_a_mirrored_periodic_domain(orc, axis, gamma):
    """LinearisedEuler reflective walls (rhs_function.c:47-176): the RHS on
    [0, N) with reflective walls along `axis` equals the RHS of the periodic
    domain [0, 2N) holding the mirror image (pressure and tangential velocity
    even, normal velocity odd) -- the wrap of the doubled domain reproduces
    the ghost rule irhs = -1-i / 2N-1-i and the sign of component 1+axis."""
    rng = np.random.default_rng(30 + axis)
    n = [10, 9, 8]
    U = rng.standard_normal((n[2], n[1], n[0], 4)) * 1e-3
    ax = 2 - axis                      # numpy axis of the grid axis
    mirror = np.flip(U, axis=ax).copy()
    mirror[..., 1 + axis] *= -1
    D = np.concatenate([U, mirror], axis=ax)
    nd = list(n)
    nd[axis] *= 2
    bc = [orc.PERIODIC] * 3
    bc[axis] = orc.REFLECTIVE
    g = orc.grid3d(tuple(n), 0.1, orc.ACOUSTIC4, gamma, bc=tuple(bc))
    gd = orc.grid3d(tuple(nd), 0.1, orc.ACOUSTIC4, gamma)
    R = orc.rhs3d(g, U)
    Rd = orc.rhs3d(gd, D)
    Rd = np.take(Rd, range(n[axis]), axis=ax)
    assert np.max(np.abs(R - Rd)) <= 1e-14 * np.max(np.abs(Rd))


@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_transmissive_walls(orc, gamma):
    """Transmissive walls copy the boundary cell: a state constant along x has
    no x-flux difference, so its RHS equals the periodic-x RHS; a constant
    state has zero RHS."""
    rng = np.random.default_rng(41)
    n = (9, 10, 8)
    col = rng.standard_normal((n[2], n[1], 1, 4)) * 1e-3
    U = np.ascontiguousarray(np.broadcast_to(col, (n[2], n[1], n[0], 4)))
    gt = orc.grid3d(n, 0.1, orc.ACOUSTIC4, gamma, bc=(orc.TRANSMISSIVE, 0, 0))
    gp = orc.grid3d(n, 0.1, orc.ACOUSTIC4, gamma)
    assert np.max(np.abs(orc.rhs3d(gt, U) - orc.rhs3d(gp, U))) <= 1e-14 * np.max(np.abs(orc.rhs3d(gp, U)))
    gall = orc.grid3d(n, 0.1, orc.ACOUSTIC4, gamma, bc=(orc.TRANSMISSIVE,) * 3)
    C0 = np.full((n[2], n[1], n[0], 4), 2.5e-4)
    assert np.max(np.abs(orc.rhs3d(gall, C0))) < 1e-15


def test_point_source_rhs_is_rhs_plus_one_delta(orc):
    """The point-source RHS (3d/weno4/point_source/rhs_function.c:337-346) is the
    plain RHS plus Source(t)/h^3 on component 1 of the source cell only, with
    Source the Ricker derivative -2 (t - t0) f0^2 exp(-f0^2 (t - t0)^2), t0 = 4/f0
    (source.c:11-17), and each call records (t, u at the receiver cell)."""
    N, h, f0 = 12, 0.5, 100.0
    g = orc.grid3d(N, h, orc.WAVE2, c0=250.0)
    U = np.random.default_rng(5).standard_normal((N, N, N, 2))
    F0 = orc.rhs3d(g, U)
    tt, tu = np.zeros(4), np.zeros(4)
    src, rcv = (3, 4, 5), (7, 8, 9)
    ps = orc.PS(g, src, rcv, f0, orc.P(tt), orc.P(tu), 0, 4)
    for q, t in enumerate((0.0373, 0.04, 0.0)):
        F = np.zeros_like(U)
        assert orc.lib().orc_rhs3d_cb_point_source(C.cast(C.pointer(ps), C.c_void_p), t, orc.P(U), orc.P(F)) == 0
        s = -2.0 * (t - 4.0 / f0) * f0 ** 2 * math.exp(-(f0 ** 2) * (t - 4.0 / f0) ** 2)
        D = F - F0
        k = (src[2], src[1], src[0], 1)
        mask = np.ones(U.shape, bool)
        mask[k] = False
        assert np.all(D[mask] == 0.0)                    # same RHS everywhere else
        # one rounding of F0 + s/h^3 at the source cell
        assert abs(D[k] - s / h ** 3) <= 4.5e-16 * (abs(F0[k]) + abs(s) / h ** 3)
        assert ps.ntrace == q + 1 and tt[q] == t and tu[q] == U[rcv[2], rcv[1], rcv[0], 0]


@pytest.mark.parametrize("system,bc", [(0, (0, 0)), (0, (1, 1)), (1, (0, 0))])
def test_2d_operators_are_the_one_plane_3d_operator(orc, system, bc):
    """With the shipped linear weights the 2D operators (WAVE2 of
    2d/weno4/standing_wave, 3-variable LLF acoustics) equal the 3D operator on
    one plane (n_z = 1, periodic z): every z-stencil of a z-invariant field
    vanishes and the face averages depend only on the 1D five-point
    combinations along the normal.  ACOUSTIC4 carries w = 0, which stays 0."""
    N, h = 12, 1.0 / 12
    nv2 = 2 if system == 0 else 3
    U2 = np.random.default_rng(3).standard_normal((N, N, nv2))
    g2 = orc.G2((N, N), h, system, 0, 1.0, 0.0, 1.0, 1.0, bc)
    R2 = np.zeros_like(U2)
    assert orc.lib().orc_rhs2d_faithful(C.byref(g2), orc.P(U2), orc.P(R2)) == 0
    U3 = np.zeros((1, N, N, 2 if system == 0 else 4))
    U3[0, ..., :nv2] = U2
    R3 = orc.rhs3d(orc.grid3d((N, N, 1), h, system, (1.0, 0.0), bc=(bc[0], bc[1], 0)), U3)
    assert np.max(np.abs(R3[0, ..., :nv2] - R2)) <= 1e-14 * np.max(np.abs(R2))
    if system == 1:
        assert np.all(R3[..., 3] == 0.0)
