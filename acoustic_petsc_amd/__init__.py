"""acoustic_petsc_amd -- MI355X (gfx950) right-hand side and SSP stepper of
the acoustic_petsc WENO-4 finite-volume solvers.

The compute path is the HIP library ``libacx.so`` built from ``csrc/`` (see
``build.py``); this module binds its C ABI (``include/acx.h``) with ctypes
and uses torch only to hold device memory and streams.  There is no CPU
fallback: every entry point raises when the library or a GPU is missing.

Mirrors the reference's operator surface:

* :func:`rhs`            -- ``RHSFunction(TS, t, U, RHS, ctx)``
                            (3d/weno4/plane_wave/rhs_function.c:12-345,
                            3d/weno3/LinearisedEuler/rhs_function.c:12-387)
* :func:`ssp_rks3_step`  -- one ``TSSSPRKS3`` step with 4 stages
                            (3d/weno4/plane_wave/hype.c:180-182)
* :class:`Grid`          -- the DMDA / AppCtx parameters the RHS reads
                            (hype.c:33-56, 102-104)
"""
import ctypes as _C
import os as _os
from dataclasses import dataclass, field

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(_HERE, "libacx.so")

WAVE2 = 0
ACOUSTIC4 = 1
PERIODIC = 0
ADIABATIC = 1        # WAVE2 walls: component 0 mirrored with its sign flipped
REFLECTIVE = 2       # ACOUSTIC4 walls: mirror, normal velocity negated
TRANSMISSIVE = 3     # ACOUSTIC4 walls: copy of the boundary cell

ACX_ERRORS = {1: "invalid argument", 2: "unsupported", 3: "HIP runtime error", 4: "non-finite state"}


class AcxError(RuntimeError):
    """A nonzero return code of the C ABI (the reference would CHKERRQ / MPI_Abort)."""

    def __init__(self, code, msg):
        super().__init__(f"acx error {code} ({ACX_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class _Grid(_C.Structure):
    _fields_ = [("n", _C.c_int32 * 3), ("system", _C.c_int32), ("bc", _C.c_int32 * 3), ("pad_", _C.c_int32),
                ("h", _C.c_double), ("gamma_hi", _C.c_double), ("gamma_lo", _C.c_double),
                ("c0", _C.c_double), ("rho0", _C.c_double)]


class _Ctcs(_C.Structure):
    _fields_ = [("n", _C.c_int32 * 3), ("stencil", _C.c_int32), ("src", _C.c_int32 * 3), ("rcv", _C.c_int32 * 3),
                ("dx", _C.c_double), ("dy", _C.c_double), ("dz", _C.c_double), ("dt", _C.c_double),
                ("f0", _C.c_double)]


@dataclass
class Grid:
    """Uniform 3D grid and operator parameters (the AppCtx of the reference).

    ``gamma_hi=1, gamma_lo=0`` is the shipped linear CWENO4
    (3d/weno4/plane_wave/reconstruction.c:42-43); ``0.85, 0.85`` the
    nonlinear setting of Viscous-Multiphase-Fourth-Order/reconstruction.c:42-43.
    """
    n: tuple
    h: float
    system: int = ACOUSTIC4
    bc: tuple = (PERIODIC, PERIODIC, PERIODIC)
    gamma_hi: float = 1.0
    gamma_lo: float = 0.0
    c0: float = 1.0
    rho0: float = 1.0

    @property
    def nvar(self):
        return 4 if self.system == ACOUSTIC4 else 2

    @property
    def ncells(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    def shape(self):
        """torch/numpy shape of a DMDA global vector: (nz, ny, nx, nvar)."""
        return (int(self.n[2]), int(self.n[1]), int(self.n[0]), self.nvar)

    def c_struct(self):
        g = _Grid()
        g.n[:] = [int(v) for v in self.n]
        g.system = int(self.system)
        g.bc[:] = [int(v) for v in self.bc]
        g.h, g.gamma_hi, g.gamma_lo, g.c0, g.rho0 = (float(self.h), float(self.gamma_hi), float(self.gamma_lo),
                                                      float(self.c0), float(self.rho0))
        return g


_lib = None


def lib():
    """The loaded libacx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m acoustic_petsc_amd.build`")
        # torch ships its own HIP runtime with the soname libamdhip64.so.7;
        # loading it first makes libacx bind to that same runtime, so device
        # pointers and streams from torch are valid in the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = _C.CDLL(LIB_PATH)
        vp, dbl, i32 = _C.c_void_p, _C.c_double, _C.c_int
        gp = _C.POINTER(_Grid)
        sigs = {
            "acx_rhs_f64": [gp, vp, vp, vp],
            "acx_stage_f64": [gp, vp, vp, vp, dbl, dbl, dbl, vp],
            "acx_ssp_rks3_step_f64": [gp, vp, vp, vp, dbl, vp],
            "acx_ssp_rks3_steps_graph_f64": [gp, vp, vp, vp, dbl, i32, vp],
            "acx_ssp_rks3_step_f32c": [gp, vp, vp, vp, vp, vp, vp, dbl, vp],
            "acx_ssp_rks3_step_f32c_sel": [gp, vp, vp, vp, vp, vp, vp, vp, dbl, vp],
            "acx_ssp_rks3_step_f32c_selm": [gp, vp, vp, vp, vp, vp, vp, vp, vp, dbl, vp],
            "acx_sel_mask": [gp, vp, vp, vp],
            "acx_norms_f64": [vp, _C.c_size_t, i32, i32, _C.POINTER(dbl), vp],
            "acx_diff_f32c_f64": [vp, vp, vp, _C.c_size_t, _C.POINTER(dbl), vp],
            "acx_vec_copy_f64": [vp, vp, _C.c_size_t, vp],
            "acx_energy_f64": [gp, vp, _C.POINTER(dbl), vp],
            "acx_ctcs_update_f64": [_C.POINTER(_Ctcs), vp, vp, vp, vp, i32, _C.POINTER(dbl), vp],
            "acx_field_gaussians": [gp, _C.c_uint64, i32, dbl, dbl, dbl, dbl, vp],
            "acx_field_plane_wave": [gp, _C.POINTER(i32), dbl, dbl, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = i32
        L.acx_last_error.restype = _C.c_char_p
        L.acx_cfl_dt.argtypes = [gp]
        L.acx_cfl_dt.argtypes = [gp, dbl]
        L.acx_cfl_dt.restype = dbl
        L.acx_sel_mask_bytes.argtypes = [gp]
        L.acx_sel_mask_bytes.restype = _C.c_size_t
        L.acx_nonfinite_check.argtypes = [vp]
        L.acx_nonfinite_check.restype = i32
        L.acx_nonfinite_clear.argtypes = []
        L.acx_nonfinite_clear.restype = i32
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise AcxError(rc, lib().acx_last_error().decode(errors="replace"))


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("acoustic_petsc_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return _C.c_void_p(s.cuda_stream)


def _ptr(t, dtype):
    torch = _torch()
    if t is None:
        return None
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"expected a contiguous CUDA tensor of {dtype}, got {t.dtype} on {t.device}")
    return _C.c_void_p(t.data_ptr())


# ---- operator surface -----------------------------------------------------

def rhs(grid, U, out=None, stream=None):
    """F(U) of the semi-discrete operator (RHSFunction), FP64 device tensors."""
    torch = _torch()
    if out is None:
        out = torch.empty_like(U)
    g = grid.c_struct()
    _check(lib().acx_rhs_f64(_C.byref(g), _ptr(U, torch.float64), _ptr(out, torch.float64), _stream(stream)))
    return out


def stage(grid, Q, U0, out, a, b, c, stream=None):
    """OUT = a*U0 + b*Q + c*F(Q), one fused SSP stage (FP64)."""
    torch = _torch()
    g = grid.c_struct()
    _check(lib().acx_stage_f64(_C.byref(g), _ptr(Q, torch.float64), _ptr(U0, torch.float64),
                               _ptr(out, torch.float64), a, b, c, _stream(stream)))
    return out


def ssp_rks3_step(grid, U, dt, work=None, stream=None):
    """One TSSSPRKS3 (4-stage) step in place on the FP64 state U."""
    torch = _torch()
    if work is None:
        work = (torch.empty_like(U), torch.empty_like(U))
    g = grid.c_struct()
    _check(lib().acx_ssp_rks3_step_f64(_C.byref(g), _ptr(U, torch.float64), _ptr(work[0], torch.float64),
                                       _ptr(work[1], torch.float64), dt, _stream(stream)))
    return U


def ssp_rks3_steps_graph(grid, U, dt, nsteps, work, stream=None):
    """nsteps TSSSPRKS3 steps in place, replayed from a captured hipGraph of
    the four stage launches (captured once per grid, arrays and dt)."""
    torch = _torch()
    g = grid.c_struct()
    _check(lib().acx_ssp_rks3_steps_graph_f64(_C.byref(g), _ptr(U, torch.float64), _ptr(work[0], torch.float64),
                                              _ptr(work[1], torch.float64), dt, int(nsteps), _stream(stream)))
    return U


class CompensatedState:
    """FP32 (hi, lo) state with compensated SSP-RK3 stepping.

    ``hi + lo`` is the value; every stage reconstructs in FP32 from
    ``hi + increment`` and the step increment is added with an error-free
    two-sum, so the state carries ~48 bits across steps while every stencil
    read is FP32.
    """

    def __init__(self, grid, U64=None):
        torch = _torch()
        self.grid = grid
        shape = grid.shape()
        self.hi = torch.zeros(shape, dtype=torch.float32, device="cuda")
        self.lo = torch.zeros_like(self.hi)
        self._hi2 = torch.empty_like(self.hi)
        self._lo2 = torch.empty_like(self.hi)
        self.d0 = torch.empty_like(self.hi)
        self.d1 = torch.empty_like(self.hi)
        if U64 is not None:
            self.hi.copy_(U64.to(torch.float32))
            self.lo.copy_((U64 - self.hi.to(torch.float64)).to(torch.float32))

    def step(self, dt, flags=None, mask=None, stream=None):
        """One step; with flags, the selective step (mask: the tile mask of
        those flags from sel_mask, else made on every call)."""
        torch = _torch()
        g = self.grid.c_struct()
        f32 = torch.float32
        if mask is not None:
            rc = lib().acx_ssp_rks3_step_f32c_selm(_C.byref(g), _ptr(self.hi, f32), _ptr(self.lo, f32),
                                                  _ptr(self._hi2, f32), _ptr(self._lo2, f32), _ptr(self.d0, f32),
                                                  _ptr(self.d1, f32), _ptr(flags, torch.uint8),
                                                  _ptr(mask, torch.uint8), dt, _stream(stream))
        elif flags is None:
            rc = lib().acx_ssp_rks3_step_f32c(_C.byref(g), _ptr(self.hi, f32), _ptr(self.lo, f32),
                                             _ptr(self._hi2, f32), _ptr(self._lo2, f32), _ptr(self.d0, f32),
                                             _ptr(self.d1, f32), dt, _stream(stream))
        else:
            rc = lib().acx_ssp_rks3_step_f32c_sel(_C.byref(g), _ptr(self.hi, f32), _ptr(self.lo, f32),
                                                 _ptr(self._hi2, f32), _ptr(self._lo2, f32), _ptr(self.d0, f32),
                                                 _ptr(self.d1, f32), _ptr(flags, torch.uint8), dt, _stream(stream))
        _check(rc)
        self.hi, self._hi2 = self._hi2, self.hi
        self.lo, self._lo2 = self._lo2, self.lo

    def value64(self):
        torch = _torch()
        return self.hi.to(torch.float64) + self.lo.to(torch.float64)


def sel_mask(grid, flags, mask=None, stream=None):
    """Tile mask of the selective step for these flags (uint8 CUDA tensor,
    reused when `mask` has the right size)."""
    torch = _torch()
    L = lib()
    g = grid.c_struct()
    n = L.acx_sel_mask_bytes(_C.byref(g))
    if n == 0:
        raise AcxError(1, "selective mask: invalid grid")
    if mask is None or mask.numel() != n:
        mask = torch.empty(n, dtype=torch.uint8, device="cuda")
    _check(L.acx_sel_mask(_C.byref(g), _ptr(flags, torch.uint8), _ptr(mask, torch.uint8), _stream(stream)))
    return mask


def norms(X, comp=-1, nvar=None, stream=None):
    """(sum of squares, max abs) with FP64 accumulation."""
    torch = _torch()
    nv = nvar if nvar is not None else X.shape[-1]
    out = (_C.c_double * 2)()
    _check(lib().acx_norms_f64(_ptr(X, torch.float64), X.numel() // nv, nv, comp, out, _stream(stream)))
    return out[0], out[1]


def vec_copy(X, Y, stream=None):
    """Y[:] = X (VecCopy) for FP64 device tensors of equal size, with the
    library's nontemporal copy kernel."""
    torch = _torch()
    if X.numel() != Y.numel():
        raise ValueError(f"vec_copy: sizes differ ({X.numel()} and {Y.numel()})")
    _check(lib().acx_vec_copy_f64(_ptr(X, torch.float64), _ptr(Y, torch.float64), X.numel(), _stream(stream)))
    return Y


def diff_f32c(hi, lo, ref, stream=None):
    """(relative L2, max abs) of (hi + lo) - ref, FP64 accumulation."""
    torch = _torch()
    out = (_C.c_double * 2)()
    _check(lib().acx_diff_f32c_f64(_ptr(hi, torch.float32), _ptr(lo, torch.float32) if lo is not None else None,
                                   _ptr(ref, torch.float64), ref.numel(), out, _stream(stream)))
    return out[0], out[1]


def energy(grid, U, stream=None):
    torch = _torch()
    out = (_C.c_double * 1)()
    g = grid.c_struct()
    _check(lib().acx_energy_f64(_C.byref(g), _ptr(U, torch.float64), out, _stream(stream)))
    return out[0]


def nonfinite_check(stream=None):
    """Synchronises the stream and raises AcxError (code 4, ACX_ENONFINITE)
    when a stage kernel on this device has written a NaN or Inf since the
    last nonfinite_clear().  RHS / stage / step calls raise the same error at
    entry once the flag is visible, without synchronising."""
    _check(lib().acx_nonfinite_check(_stream(stream)))


def nonfinite_clear():
    """Resets the non-finite watch of the current device."""
    _check(lib().acx_nonfinite_clear())


def cfl_dt(grid, cfl):
    g = grid.c_struct()
    return lib().acx_cfl_dt(_C.byref(g), cfl)


# ---- CTCS sibling -------------------------------------------------------

@dataclass
class Ctcs:
    """3d/ctcs/point_source/wave.c:58-81 parameters."""
    n: tuple = (100, 100, 100)
    dx: float = 0.5
    dy: float = 0.5
    dz: float = 0.5
    dt: float = 0.0005
    c0: float = 250.0
    stencil: int = 2
    src: tuple = (50, 50, 50)
    rcv: tuple = (57, 57, 57)
    f0: float = 100.0

    def c_struct(self):
        c = _Ctcs()
        c.n[:] = list(self.n)
        c.stencil = self.stencil
        c.src[:] = list(self.src)
        c.rcv[:] = list(self.rcv)
        c.dx, c.dy, c.dz, c.dt, c.f0 = self.dx, self.dy, self.dz, self.dt, self.f0
        return c


def ctcs_run(p, nt, stream=None):
    """UpdatePressure loop of wave.c:109-121 on the GPU; returns the receiver trace."""
    torch = _torch()
    n = (p.n[2], p.n[1], p.n[0])
    Pold = torch.zeros(n, dtype=torch.float64, device="cuda")
    P = torch.zeros_like(Pold)
    Pnew = torch.zeros_like(Pold)
    C = torch.full_like(Pold, p.c0)
    cs = p.c_struct()
    trace = []
    r = (_C.c_double * 1)()
    for it in range(nt):
        _check(lib().acx_ctcs_update_f64(_C.byref(cs), _ptr(Pold, torch.float64), _ptr(P, torch.float64),
                                         _ptr(Pnew, torch.float64), _ptr(C, torch.float64), it, r, _stream(stream)))
        trace.append(r[0])
        Pold, P, Pnew = P, Pnew, Pold   # Pold = P; P = Pnew (VecCopy pair, wave.c:115-116)
    return trace


# ---- synthetic fields (host) ---------------------------------------------

def field_gaussians(grid, seed, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08)):
    """Sum of seeded Gaussian pressure pulses (exact cell averages), numpy (nz,ny,nx,nvar)."""
    import numpy as np
    U = np.zeros(grid.shape(), dtype=np.float64)
    g = grid.c_struct()
    _check(lib().acx_field_gaussians(_C.byref(g), seed, npulse, amp[0], amp[1], sigma[0], sigma[1],
                                     U.ctypes.data_as(_C.c_void_p)))
    return U


def field_plane_wave(grid, k, A=1e-3, t=0.0):
    """Standing acoustic plane wave (exact cell averages), numpy (nz,ny,nx,4)."""
    import numpy as np
    U = np.zeros(grid.shape(), dtype=np.float64)
    g = grid.c_struct()
    kk = (_C.c_int32 * 3)(*[int(v) for v in k])
    _check(lib().acx_field_plane_wave(_C.byref(g), kk, A, t, U.ctypes.data_as(_C.c_void_p)))
    return U


def field_noise(grid, seed, kmin=4.0, kmax=16.0, rms=1e-4, pulse_amp=1e-3, pulse_sigma=0.05):
    """Band-limited random-phase pressure noise (integer k, kmin <= |k| <= kmax,
    exact cell averages through the sinc factors) scaled to RMS `rms`, plus one
    seeded Gaussian pulse; zero velocity.  numpy (nz,ny,nx,4)."""
    import numpy as np
    nz, ny, nx, nv = grid.shape()
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.fft.rfftfreq(nz, 1.0 / nz)
    KZ, KY, KX = kz[None, None, :], ky[None, :, None], kx[:, None, None]
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    phase = rng.uniform(0, 2 * np.pi, size=band.shape)
    sinc = np.sinc(KX / nx) * np.sinc(KY / ny) * np.sinc(KZ / nz)
    spec = np.where(band, np.exp(1j * phase) * sinc, 0.0)
    # axes of spec are (x, y, z-half); inverse real FFT over z last
    p = np.fft.irfftn(spec, s=(nx, ny, nz), axes=(0, 1, 2))
    p = np.transpose(p, (2, 1, 0))
    p *= rms / np.sqrt(np.mean(p ** 2))
    U = field_gaussians(grid, seed + 1, npulse=1, amp=(pulse_amp, pulse_amp), sigma=(pulse_sigma, pulse_sigma))
    U[..., 0] += p
    return U
