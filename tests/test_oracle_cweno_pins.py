"""Pins of the oracle's nonlinear CWENO4 on data where the answer is known.

The sided and central quadratic stencils of the CWENO4 reconstruction
(3d/weno4/plane_wave/reconstruction.c:121-203; 2D:
2d/weno4/bubble/reconstruction.c:161-199) each read a fixed set of cells: a
sided stencil only the half-space on its side of the cell (e.g. -x: the
cells with x <= 0), the central one only the 3x3(x3) block around the cell.
Each is exact for quadratics.  Two consequences fix every derivative
estimate and the nonlinear combination (reconstruction.c:205-257; 2D
:203-249) without restating a formula:

1. Footprint: a block that holds the cell averages of a quadratic q on a
   stencil's footprint and arbitrary values elsewhere must give that
   stencil's 9 (2D: 5) derivative estimates equal to q's coefficients.
2. Selection: when every other stencil (and the 4th-order one) reads a jump
   of size J >> |q| while one stencil reads only q, the nonlinear weights
   give that stencil all the mass, so the reconstructed coefficients must
   equal q's (the first and second derivatives) and the cubic ones vanish.

Both are checked on the real oracle and on mutants of it compiled here from
oracle/cweno.c with one planted mistake each (the weights gamma in place of
w in the nonlinear sum, the derivative formulas of two opposite sided
stencils swapped, a diagonal taken from the wrong side); every mutant must
fail them.  The mutants are built on the CPU only.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
dp = C.POINTER(C.c_double)

GAMMA = (0.85, 0.85)        # the reference's nonlinear setting (Viscous-Multiphase-Fourth-Order)
JUMP = 1.0e3

# 3D stencil order of the oracle: central, +x, -x, +y, -y, +z, -z
STENCILS_3D = [None, (2, +1), (2, -1), (1, +1), (1, -1), (0, +1), (0, -1)]   # (block axis [z,y,x], side)
# 2D stencil order: central, -x, +x, -y, +y
STENCILS_2D = [None, (1, -1), (1, +1), (0, -1), (0, +1)]                      # (block axis [y,x], side)


def _bind(lib):
    lib.orc_cweno3d_ext.argtypes = [dp, C.c_ssize_t, C.c_ssize_t, C.c_ssize_t, C.c_double, C.c_double,
                                    dp, dp, dp]
    lib.orc_cweno2d_ext.argtypes = [dp, C.c_ssize_t, C.c_ssize_t, C.c_double, C.c_double, dp, dp, dp]
    return lib


def _p(a):
    return a.ctypes.data_as(dp)


def cweno3d_ext(lib, blk, gamma=GAMMA):
    blk = np.ascontiguousarray(blk, np.float64)
    co, dg, ds = np.zeros(20), np.zeros(10), np.zeros(63)
    centre = C.cast(blk.ctypes.data + 8 * (2 * 25 + 2 * 5 + 2), dp)
    lib.orc_cweno3d_ext(centre, 1, 5, 25, gamma[0], gamma[1], _p(co), _p(dg), _p(ds))
    return co, dg, ds.reshape(7, 9)


def cweno2d_ext(lib, blk, gamma=GAMMA):
    blk = np.ascontiguousarray(blk, np.float64)
    co, dg, ds = np.zeros(10), np.zeros(8), np.zeros(25)
    centre = C.cast(blk.ctypes.data + 8 * (2 * 5 + 2), dp)
    lib.orc_cweno2d_ext(centre, 1, 5, gamma[0], gamma[1], _p(co), _p(dg), _p(ds))
    return co, dg, ds.reshape(5, 5)


# --- quadratics and their exact cell averages ------------------------------
#
# In the reconstruction's basis (reconstruction.c:12-21) the quadratic part is
# c0 + c1 x + c2 y + c3 z + c4 (x^2 - 1/12) + c5 (y^2 - 1/12) + c6 (z^2 - 1/12)
# + c7 xy + c8 yz + c9 zx on cells of unit width.  Over the unit cell centred
# at the integer point (i, j, k) the average of x^2 - 1/12 is i^2 and that of
# xy is ij, so the cell averages are the polynomial's coefficients applied to
# the monomials of the cell centre.

def quad3d_averages(c):
    k, j, i = np.meshgrid(np.arange(-2, 3), np.arange(-2, 3), np.arange(-2, 3), indexing="ij")
    return (c[0] + c[1] * i + c[2] * j + c[3] * k + c[4] * i * i + c[5] * j * j + c[6] * k * k
            + c[7] * i * j + c[8] * j * k + c[9] * k * i).astype(np.float64)


def quad2d_averages(c):
    # 2D basis (2d/weno4/bubble): c0 + c1 x + c2 y + c3 (x^2-1/12) + c4 (y^2-1/12) + c5 xy
    j, i = np.meshgrid(np.arange(-2, 3), np.arange(-2, 3), indexing="ij")
    return (c[0] + c[1] * i + c[2] * j + c[3] * i * i + c[4] * j * j + c[5] * i * j).astype(np.float64)


def footprint(shape_dims, stencil):
    """Boolean mask of the cells a stencil may read: the 3x3(x3) block for the
    central stencil, the closed half-space on its side for a sided one."""
    idx = np.meshgrid(*[np.arange(-2, 3)] * shape_dims, indexing="ij")
    if stencil is None:
        m = np.ones(idx[0].shape, bool)
        for a in idx:
            m &= np.abs(a) <= 1
        return m
    ax, side = stencil
    return side * idx[ax] >= 0


def jump_outside(shape_dims, stencil, rng):
    """Values J (1 + r) with r in [0, 1), signed so that every stencil that
    reaches outside the footprint sees a first difference of order J: the
    far half-space of a sided stencil gets +J (1 + r); for the central
    stencil the +2 end of every axis arm gets -J (1 + r), the -2 end +J (1 + r)."""
    idx = np.meshgrid(*[np.arange(-2, 3)] * shape_dims, indexing="ij")
    mag = JUMP * (1.0 + rng.random(idx[0].shape))
    if stencil is None:
        sign = -np.sign(sum(idx))
        sign[sign == 0] = 1.0
        return sign * mag
    return mag


# --- the two checks, for any build of the oracle ----------------------------

def footprint_errors_3d(lib, seed=0):
    """max |stencil estimate - exact| per stencil, junk outside its footprint."""
    rng = np.random.default_rng(seed)
    errs = []
    for s, st in enumerate(STENCILS_3D):
        c = np.zeros(10)
        c[:] = rng.standard_normal(10)
        blk = quad3d_averages(c)
        m = footprint(3, st)
        blk[~m] = JUMP * rng.standard_normal(int((~m).sum()))
        _, _, ds = cweno3d_ext(lib, blk)
        errs.append(float(np.max(np.abs(ds[s] - c[1:10]))))
    return errs


def selection_errors_3d(lib, seed=1):
    """(max |coeff - exact|, 1 - w of the clean stencil) per stencil."""
    rng = np.random.default_rng(seed)
    out = []
    for s, st in enumerate(STENCILS_3D):
        c = rng.standard_normal(10)
        blk = quad3d_averages(c)
        m = footprint(3, st)
        blk[~m] += jump_outside(3, st, rng)[~m]
        co, dg, _ = cweno3d_ext(lib, blk)
        want = np.zeros(20)
        want[:10] = c
        want[0] = blk[2, 2, 2]
        out.append((float(np.max(np.abs(co - want))), float(1.0 - dg[2 + s])))
    return out


def footprint_errors_2d(lib, seed=2):
    rng = np.random.default_rng(seed)
    errs = []
    for s, st in enumerate(STENCILS_2D):
        c = rng.standard_normal(6)
        blk = quad2d_averages(c)
        m = footprint(2, st)
        blk[~m] = JUMP * rng.standard_normal(int((~m).sum()))
        _, _, ds = cweno2d_ext(lib, blk)
        errs.append(float(np.max(np.abs(ds[s] - c[1:6]))))
    return errs


def selection_errors_2d(lib, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for s, st in enumerate(STENCILS_2D):
        c = rng.standard_normal(6)
        blk = quad2d_averages(c)
        m = footprint(2, st)
        blk[~m] += jump_outside(2, st, rng)[~m]
        co, dg, _ = cweno2d_ext(lib, blk)
        want = np.zeros(10)
        want[:6] = c
        want[0] = blk[2, 2]
        out.append((float(np.max(np.abs(co - want))), float(1.0 - dg[2 + s])))
    return out


TOL = 1e-11


def pins_hold(lib):
    """True when every footprint and selection check passes."""
    if max(footprint_errors_3d(lib)) > TOL or max(footprint_errors_2d(lib)) > TOL:
        return False
    for err, miss in selection_errors_3d(lib) + selection_errors_2d(lib):
        if err > TOL or miss > 1e-12:
            return False
    return True


# --- tests on the oracle ----------------------------------------------------

@pytest.fixture(scope="module")
def olib(orc):
    return _bind(orc.lib())


def test_quadratic_cell_averages_closed_form():
    """The closed-form cell averages used here agree with Gauss quadrature of
    the basis polynomials (3-point rule, exact for quadratics)."""
    gx, gw = np.polynomial.legendre.leggauss(3)
    gx, gw = gx / 2, gw / 2
    c = np.random.default_rng(7).standard_normal(10)

    def q(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * (x * x - 1 / 12) + c[5] * (y * y - 1 / 12)
                + c[6] * (z * z - 1 / 12) + c[7] * x * y + c[8] * y * z + c[9] * z * x)

    blk = quad3d_averages(c)
    for (k, j, i) in [(0, 0, 0), (1, -2, 2), (-1, 1, -2), (2, 2, 2)]:
        A, B, Cc = np.meshgrid(i + gx, j + gx, k + gx, indexing="ij")
        W = gw[:, None, None] * gw[None, :, None] * gw[None, None, :]
        assert abs(np.sum(W * q(A, B, Cc)) - blk[k + 2, j + 2, i + 2]) < 1e-12


@pytest.mark.parametrize("seed", [0, 10, 20])
def test_every_3d_stencil_reads_only_its_footprint(olib, seed):
    errs = footprint_errors_3d(olib, seed)
    assert max(errs) < TOL, errs


@pytest.mark.parametrize("seed", [1, 11, 21])
def test_3d_weights_select_the_stencil_that_avoids_the_jump(olib, seed):
    for s, (err, miss) in enumerate(selection_errors_3d(olib, seed)):
        assert miss < 1e-12, (s, miss)
        assert err < TOL, (s, err)


@pytest.mark.parametrize("seed", [2, 12])
def test_every_2d_stencil_reads_only_its_footprint(olib, seed):
    errs = footprint_errors_2d(olib, seed)
    assert max(errs) < TOL, errs


@pytest.mark.parametrize("seed", [3, 13])
def test_2d_weights_select_the_stencil_that_avoids_the_jump(olib, seed):
    for s, (err, miss) in enumerate(selection_errors_2d(olib, seed)):
        assert miss < 1e-12, (s, miss)
        assert err < TOL, (s, err)


# --- planted mistakes (mutants of oracle/cweno.c, CPU only) -------------------

MUTANTS = {
    "gamma_in_nonlinear_sum_3d": [("nl[q] += w3[s] * d[s][q];", "nl[q] += g3[s] * d[s][q];")],
    "swap_xy_of_the_x_stencils_3d": [
        ("d[1][DXY] = 0.5 * (xy_pp - xy_pm) - d[1][DY];", "d[1][DXY] = 0.5 * (xy_mm - xy_mp) + d[1][DY];"),
        ("d[2][DXY] = 0.5 * (xy_mm - xy_mp) + d[2][DY];", "d[2][DXY] = 0.5 * (xy_pp - xy_pm) - d[2][DY];")],
    "swap_yz_of_the_y_stencils_3d": [
        ("d[3][DYZ] = 0.5 * (yz_pp - yz_pm) - d[3][DZ];", "d[3][DYZ] = 0.5 * (yz_mm - yz_mp) + d[3][DZ];"),
        ("d[4][DYZ] = 0.5 * (yz_mm - yz_mp) + d[4][DZ];", "d[4][DYZ] = 0.5 * (yz_pp - yz_pm) - d[4][DZ];")],
    "plus_z_xz_from_the_wrong_diagonal_3d": [
        ("d[5][DXZ] = 0.5 * (zx_pp - zx_pm) - d[5][DX];", "d[5][DXZ] = 0.5 * (zx_pp - zx_mp) - d[5][DX];")],
    "minus_z_yz_from_the_wrong_diagonal_3d": [
        ("d[6][DYZ] = 0.5 * (yz_mm - yz_pm) + d[6][DY];", "d[6][DYZ] = 0.5 * (yz_mm - yz_mp) + d[6][DY];")],
    "gamma_in_nonlinear_sum_2d": [
        ("w3[0] * D[0] + w3[1] * D[1] + w3[2] * D[2] + w3[3] * D[3] + w3[4] * D[4]",
         "g3[0] * D[0] + g3[1] * D[1] + g3[2] * D[2] + g3[3] * D[3] + g3[4] * D[4]")],
    "swap_xy_of_the_x_stencils_2d": [
        ("dxy[1] = 0.5 * (mm - mp - ym1 + yp1);", "dxy[1] = 0.5 * (-pm + pp + ym1 - yp1);"),
        ("dxy[2] = 0.5 * (-pm + pp + ym1 - yp1);", "dxy[2] = 0.5 * (mm - mp - ym1 + yp1);")],
}


def _build_mutant(tmp_path, name, edits):
    with open(os.path.join(ORACLE_DIR, "cweno.c")) as f:
        src = f.read()
    for old, new in edits:
        assert src.count(old) == 1, f"{name}: the line to mutate is not unique or missing: {old}"
        src = src.replace(old, new)
    path = tmp_path / f"{name}.c"
    path.write_text(src)
    so = tmp_path / f"lib{name}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c11", f"-I{ORACLE_DIR}",
                    str(path), "-o", str(so), "-lm"], check=True)
    return _bind(C.CDLL(str(so)))


def test_unmutated_copy_passes(tmp_path):
    """The compile-and-load route of the mutants, with no edit, passes the pins."""
    assert pins_hold(_build_mutant(tmp_path, "identity", []))


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(tmp_path, name):
    lib = _build_mutant(tmp_path, name, MUTANTS[name])
    assert not pins_hold(lib), f"planted mistake {name} passed the CWENO4 pins"
