"""The C ABI library loads and exports every function declared in include/*.h
(no GPU calls: only symbol resolution and host-side helpers)."""
import ctypes as C
import glob
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"^\s*(typedef|#).*$", "", text, flags=re.M)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "acx_rhs_f64" in names and "acx_ssp_rks3_step_f64" in names
    assert "RHSFunction" in names and "TSSolve" in names and "DMDACreate3d" in names


def test_library_exports_every_declared_symbol(acx):
    lib = acx.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_field_helpers(acx):
    g = acx.Grid(n=(16, 16, 16), h=1.0 / 16)
    U = acx.field_gaussians(g, 7, npulse=2, amp=(1e-3, 1e-3), sigma=(0.1, 0.1))
    assert U.shape == (16, 16, 16, 4)
    assert np.all(U[..., 1:] == 0) and U[..., 0].max() > 0
    # the cell averages of one Gaussian integrate to its total mass
    U1 = acx.field_gaussians(g, 7, npulse=1, amp=(1.0, 1.0), sigma=(0.1, 0.1))
    mass = U1[..., 0].sum() / 16 ** 3
    assert abs(mass - (2 * np.pi) ** 1.5 * 0.1 ** 3) < 1e-12
    P = acx.field_plane_wave(g, (1, 2, 1), A=1e-3)
    assert abs(P[..., 0].mean()) < 1e-18 and np.all(P[..., 1:] == 0)


def test_invalid_arguments_are_reported(acx):
    """Argument checks run on the host and return an error code, no GPU needed."""
    lib = acx.lib()
    g = acx.Grid(n=(2, 16, 16), h=0.1).c_struct()
    rc = lib.acx_rhs_f64(C.byref(g), C.c_void_p(16), C.c_void_p(32), None)
    assert rc == 1 and b"at least 4 cells" in lib.acx_last_error()
    g = acx.Grid(n=(16, 16, 16), h=0.1, bc=(acx.ADIABATIC, 0, 0)).c_struct()   # ACOUSTIC4 default
    rc = lib.acx_rhs_f64(C.byref(g), C.c_void_p(16), C.c_void_p(32), None)
    assert rc == 2 and b"use reflective" in lib.acx_last_error()
    g = acx.Grid(n=(16, 16, 16), h=0.1, system=acx.WAVE2, bc=(0, acx.TRANSMISSIVE, 0)).c_struct()
    rc = lib.acx_rhs_f64(C.byref(g), C.c_void_p(16), C.c_void_p(32), None)
    assert rc == 2 and b"ACOUSTIC4 rules" in lib.acx_last_error()
    g = acx.Grid(n=(16, 16, 16), h=0.1, bc=(0, 0, 7)).c_struct()
    rc = lib.acx_rhs_f64(C.byref(g), C.c_void_p(16), C.c_void_p(32), None)
    assert rc == 1 and b"not a boundary kind" in lib.acx_last_error()
    g = acx.Grid(n=(16, 16, 16), h=0.1).c_struct()
    rc = lib.acx_stage_f64(C.byref(g), C.c_void_p(16), None, C.c_void_p(16), 0.0, 1.0, 1.0, None)
    assert rc == 1


def test_compensated_step_rejects_aliased_arrays(acx):
    """The six arrays of the FP32 compensated step (and of the selective
    step) must be pairwise distinct: the 2.5D march reads HI and the
    increments of neighbouring tiles as halo while other tiles write."""
    lib = acx.lib()
    g = acx.Grid(n=(16, 16, 16), h=0.1).c_struct()
    base = [C.c_void_p(0x1000 * (k + 1)) for k in range(6)]
    flags = C.c_void_p(0x9000)
    for i in range(6):
        for j in range(i):
            arr = list(base)
            arr[i] = arr[j]
            rc = lib.acx_ssp_rks3_step_f32c(C.byref(g), *arr, 1e-3, None)
            assert rc == 1 and b"must be distinct" in lib.acx_last_error(), (i, j)
            rc = lib.acx_ssp_rks3_step_f32c_sel(C.byref(g), *arr, flags, 1e-3, None)
            assert rc == 1 and b"must be distinct" in lib.acx_last_error(), (i, j)
