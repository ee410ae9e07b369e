"""ctypes view of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

WAVE2, ACOUSTIC4 = 0, 1
PERIODIC, ADIABATIC, REFLECTIVE, TRANSMISSIVE = 0, 1, 2, 3

dp = C.POINTER(C.c_double)


class G3(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("h", C.c_double), ("system", C.c_int), ("gamma_hi", C.c_double),
                ("gamma_lo", C.c_double), ("c0", C.c_double), ("rho0", C.c_double), ("cfl", C.c_double),
                ("bc", C.c_int * 3)]


class G2(C.Structure):
    _fields_ = [("N", C.c_int * 2), ("h", C.c_double), ("system", C.c_int), ("nonlinear", C.c_int),
                ("gamma_hi", C.c_double), ("gamma_lo", C.c_double), ("c0", C.c_double), ("rho0", C.c_double),
                ("bc", C.c_int * 2)]


class CT(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("c0", C.c_double), ("stencil", C.c_int), ("src", C.c_int * 3), ("rcv", C.c_int * 3),
                ("f0", C.c_double)]


class PS(C.Structure):
    _fields_ = [("g", G3), ("src", C.c_int * 3), ("rcv", C.c_int * 3), ("f0", C.c_double), ("trace_t", dp),
                ("trace_u", dp), ("ntrace", C.c_int), ("cap", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_ts_solve.restype = C.c_int
        L.orc_ts_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, dp, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_int]
        L.orc_ssp_rks3_step.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, dp, C.c_double, C.c_double, C.c_int,
                                        dp, dp, dp]
        L.orc_ssp_rk104_step.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, dp, C.c_double, C.c_double, dp, dp, dp]
        L.orc_rhs3d_faithful.argtypes = [C.POINTER(G3), dp, dp, dp]
        L.orc_rhs3d_stream.argtypes = [C.POINTER(G3), dp, dp, C.c_int]
        L.orc_rhs2d_faithful.argtypes = [C.POINTER(G2), dp, dp]
        L.orc_cweno3d.argtypes = [dp, C.c_ssize_t, C.c_ssize_t, C.c_ssize_t, C.c_double, C.c_double, dp, dp]
        L.orc_cweno2d.argtypes = [dp, C.c_ssize_t, C.c_ssize_t, C.c_double, C.c_double, dp]
        L.orc_weno2d_linear.argtypes = [dp, C.c_ssize_t, C.c_ssize_t, dp]
        L.orc_poly3d.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_poly3d.restype = C.c_double
        L.orc_grad3d.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.orc_plane_wave_init.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.orc_plane_wave_error.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, dp]
        L.orc_standing_wave2d_init.argtypes = [C.c_int, dp]
        L.orc_standing_wave2d_error.argtypes = [C.c_int, C.c_double, dp, dp, dp]
        L.orc_ctcs3d_run.argtypes = [C.POINTER(CT), C.c_int, dp]
        L.orc_ctcs3d_update.argtypes = [C.POINTER(CT), dp, dp, dp, dp, C.c_int]
        L.orc_ctcs3d_update.restype = C.c_double
        L.orc_rhs3d_cb_point_source.argtypes = [C.c_void_p, C.c_double, dp, dp]
        L.orc_weno_source.argtypes = [C.c_double, C.c_double]
        L.orc_weno_source.restype = C.c_double
        _lib = L
    return _lib


def P(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def grid3d(N, h, system=ACOUSTIC4, gamma=(1.0, 0.0), c0=1.0, rho0=1.0, cfl=0.4, bc=(0, 0, 0)):
    n = (N, N, N) if np.isscalar(N) else tuple(N)
    return G3(n, h, system, gamma[0], gamma[1], c0, rho0, cfl, tuple(bc))


def nvar(system):
    return 2 if system == WAVE2 else 4


def rhs3d(g, U, faithful=True):
    """RHS of a DMDA global vector (numpy (nz,ny,nx,nvar))."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    R = np.zeros_like(U)
    dt = C.c_double()
    rc = lib().orc_rhs3d_faithful(C.byref(g), P(U), P(R), C.byref(dt)) if faithful else \
        lib().orc_rhs3d_stream(C.byref(g), P(U), P(R), 0)
    if rc:
        raise RuntimeError(f"oracle rhs failed ({rc})")
    return R


def cweno3d(block5, gamma=(1.0, 0.0)):
    """CWENO4 coefficients of the centre of a 5x5x5 block indexed [z][y][x]; returns (coeffs, diag)."""
    b = np.ascontiguousarray(block5, dtype=np.float64)
    assert b.shape == (5, 5, 5)
    co = np.zeros(20)
    dg = np.zeros(10)
    centre = b.ctypes.data + 8 * (2 * 25 + 2 * 5 + 2)
    lib().orc_cweno3d(C.cast(centre, dp), 1, 5, 25, gamma[0], gamma[1], P(co), P(dg))
    return co, dg


def poly3d(x, y, z, co):
    return lib().orc_poly3d(x, y, z, P(np.ascontiguousarray(co, dtype=np.float64)))


def ts_solve(cb_name, ctx, U, dt, t_final, scheme=0, stages=4, t0=0.0):
    """orc_ts_solve with one of the oracle's RHS callbacks (by symbol name)."""
    fn = C.cast(getattr(lib(), cb_name), C.c_void_p)
    steps = lib().orc_ts_solve(fn, C.byref(ctx), U.size, P(U), t0, dt, t_final, scheme, stages)
    if steps < 0:
        raise RuntimeError("oracle ts_solve failed")
    return steps


def plane_wave_init(N, h, t=0.0):
    U = np.zeros((N, N, N, 2))
    lib().orc_plane_wave_init(N, h, t, P(U))
    return U


def plane_wave_error(N, h, t, U):
    l2, li = C.c_double(), C.c_double()
    lib().orc_plane_wave_error(N, h, t, P(U), C.byref(l2), C.byref(li))
    return l2.value, li.value


def ctcs_trace(n=(100, 100, 100), dx=0.5, dt=0.0005, c0=250.0, stencil=2, src=(50, 50, 50), rcv=(57, 57, 57),
               f0=100.0, nt=251):
    c = CT(n, dx, dx, dx, dt, c0, stencil, src, rcv, f0)
    tr = np.zeros(nt)
    lib().orc_ctcs3d_run(C.byref(c), nt, P(tr))
    return tr


def ctcs_steps(n, d, dt, stencil, src, rcv, Pold, P0, Cf, nt, it0=0, f0=100.0):
    """nt leapfrog updates (wave.c:109-121) from the pair (Pold, P0), per-cell speed
    Cf (numpy (nz,ny,nx)); returns (Pold, P, receiver trace)."""
    c = CT(tuple(n), d[0], d[1], d[2], dt, 0.0, stencil, tuple(src), tuple(rcv), f0)
    a, b = np.array(Pold, np.float64), np.array(P0, np.float64)
    new = np.zeros_like(b)
    Cf = np.ascontiguousarray(Cf, np.float64)
    tr = np.zeros(nt)
    for it in range(nt):
        tr[it] = lib().orc_ctcs3d_update(C.byref(c), P(a), P(b), P(new), P(Cf), it0 + it)
        a, b, new = b, new, a
    return a, b, tr


def point_source_run(N, h, c0, dt, nt, src, rcv, f0=100.0, gamma=(1.0, 0.0)):
    """The WENO point-source program (3d/weno4/point_source/hype.c:14-183): WAVE2
    from rest, one direct RHS call at t = 0 (hype.c:170), then nt SSP-RK3 (s = 4)
    steps.  Returns (state (N,N,N,2), receiver times, receiver values): one
    receiver sample per RHS evaluation, as rhs_function.c:342-346 writes them."""
    g = grid3d(N, h, WAVE2, gamma, c0=c0)
    U = np.zeros((N, N, N, 2))
    cap = 4 * nt + 16
    tt, tu = np.zeros(cap), np.zeros(cap)
    ps = PS(g, tuple(src), tuple(rcv), f0, P(tt), P(tu), 0, cap)
    F = np.zeros_like(U)
    if lib().orc_rhs3d_cb_point_source(C.cast(C.pointer(ps), C.c_void_p), 0.0, P(U), P(F)):
        raise RuntimeError("oracle point-source rhs failed")
    ts_solve("orc_rhs3d_cb_point_source", ps, U, dt, nt * dt, 0, 4)
    return U, tt[:ps.ntrace].copy(), tu[:ps.ntrace].copy()
