/*
 * cweno.c -- oracle restatement of the CWENO4 reconstructions (test
 * infrastructure only; see oracle.h).
 *
 *   3D: 3d/weno4/plane_wave/reconstruction.c:12-259 (gamma values passed in;
 *       the shipped plane_wave file hard-codes (1,0) at :42-43, the nonlinear
 *       setting (0.85,0.85) is Viscous-Multiphase-Fourth-Order/reconstruction.c:42-43)
 *   2D: 2d/weno4/standing_wave/reconstruction.c:106-125 (linear) and
 *       2d/weno4/bubble/reconstruction.c:123-251 (CWENO4)
 */
#include <math.h>
#include "oracle.h"

/* constants of hype.h:37,86-93 */
static const double k_eps   = 1.0e-12;
static const double k_1_3   = 1. / 3.;
static const double k_1_6   = 1. / 6.;
static const double k_1_12  = 1. / 12.;
static const double k_3_20  = 3. / 20.;
static const double k_1_120 = 1. / 120.;
static const double k_13_3  = 13. / 3.;
static const double k_7_6   = 7. / 6.;
static const double k_61_48 = 61. / 48.;
static const double k_11_120 = 11. / 120.;   /* 2d/weno4/standing_wave/hype.h:97 */
static const double k_41_60  = 41. / 60.;    /* 2d/weno4/standing_wave/hype.h:99 */

/* The 9 derivative estimates of one quadratic (R3) stencil, in the order
 * x, y, z, xx, yy, zz, xy, yz, xz. */
enum { DX, DY, DZ, DXX, DYY, DZZ, DXY, DYZ, DXZ, NQ };

static double smooth_q(const double d[NQ])
{
    /* IS of a quadratic: sum over 1<=|alpha|<=2 of the cell integral of
     * (D^alpha p)^2 (reconstruction.c:206-208) */
    return (d[DX] * d[DX] + d[DY] * d[DY] + d[DZ] * d[DZ]) +
           k_13_3 * (d[DXX] * d[DXX] + d[DYY] * d[DYY] + d[DZZ] * d[DZZ]) +
           k_7_6 * (d[DXY] * d[DXY] + d[DYZ] * d[DYZ] + d[DXZ] * d[DXZ]);
}

void orc_cweno3d(const double *c, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                 double gamma_hi, double gamma_lo, double co[20], double *diag)
{
    orc_cweno3d_ext(c, sx, sy, sz, gamma_hi, gamma_lo, co, diag, NULL);
}

void orc_cweno3d_ext(const double *c, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                     double gamma_hi, double gamma_lo, double co[20], double *diag, double *dstencil)
{
#define V(a, b, e) c[(a) * sx + (b) * sy + (e) * sz]
    const double u0 = V(0, 0, 0);
    /* axis arms */
    const double xm2 = V(-2, 0, 0), xm1 = V(-1, 0, 0), xp1 = V(1, 0, 0), xp2 = V(2, 0, 0);
    const double ym2 = V(0, -2, 0), ym1 = V(0, -1, 0), yp1 = V(0, 1, 0), yp2 = V(0, 2, 0);
    const double zm2 = V(0, 0, -2), zm1 = V(0, 0, -1), zp1 = V(0, 0, 1), zp2 = V(0, 0, 2);
    /* planar diagonals, named by the signs of (first axis, second axis) */
    const double xy_pp = V(1, 1, 0), xy_pm = V(1, -1, 0), xy_mp = V(-1, 1, 0), xy_mm = V(-1, -1, 0);
    const double yz_pp = V(0, 1, 1), yz_pm = V(0, 1, -1), yz_mp = V(0, -1, 1), yz_mm = V(0, -1, -1);
    const double zx_pp = V(1, 0, 1), zx_pm = V(-1, 0, 1), zx_mp = V(1, 0, -1), zx_mm = V(-1, 0, -1);
    /* corners, signs in (x,y,z) */
    const double ppp = V(1, 1, 1), mpp = V(-1, 1, 1), pmp = V(1, -1, 1), ppm = V(1, 1, -1);
    const double mmp = V(-1, -1, 1), mpm = V(-1, 1, -1), pmm = V(1, -1, -1), mmm = V(-1, -1, -1);
#undef V

    /* linear weights (reconstruction.c:55-59) */
    const double g4 = gamma_hi;
    double g3[7];
    g3[0] = (1.0 - gamma_hi) * gamma_lo;
    for (int s = 1; s < 7; ++s) g3[s] = k_1_6 * (1 - gamma_hi) * (1 - gamma_lo);

    /* fourth-order central polynomial (reconstruction.c:90-108) */
    const double ux = k_1_120 * (11.0 * xm2 - 82.0 * xm1 + 82.0 * xp1 - 11.0 * xp2);
    const double uy = k_1_120 * (11.0 * ym2 - 82.0 * ym1 + 82.0 * yp1 - 11.0 * yp2);
    const double uz = k_1_120 * (11.0 * zm2 - 82.0 * zm1 + 82.0 * zp1 - 11.0 * zp2);
    const double uxx = 0.5 * (xm1 - 2.0 * u0 + xp1);
    const double uyy = 0.5 * (ym1 - 2.0 * u0 + yp1);
    const double uzz = 0.5 * (zm1 - 2.0 * u0 + zp1);
    const double uxy = 0.25 * (xy_mm - xy_mp - xy_pm + xy_pp);
    const double uyz = 0.25 * (yz_mm - yz_mp - yz_pm + yz_pp);
    const double uxz = 0.25 * (zx_mm - zx_mp - zx_pm + zx_pp);
    const double uxxx = k_1_12 * (-xm2 + 2.0 * (xm1 - xp1) + xp2);
    const double uyyy = k_1_12 * (-ym2 + 2.0 * (ym1 - yp1) + yp2);
    const double uzzz = k_1_12 * (-zm2 + 2.0 * (zm1 - zp1) + zp2);
    const double uxxy = 0.25 * (-xy_mm + xy_mp - xy_pm + xy_pp) + 0.5 * (ym1 - yp1);
    const double uxyy = 0.5 * (xm1 - xp1) + 0.25 * (xy_pp - xy_mm - xy_mp + xy_pm);
    const double uyyz = 0.25 * (-yz_mm + yz_mp - yz_pm + yz_pp) + 0.5 * (zm1 - zp1);
    const double uyzz = 0.5 * (ym1 - yp1) + 0.25 * (yz_pp - yz_mm - yz_mp + yz_pm);
    const double uxxz = 0.25 * (-zx_mm + zx_pm - zx_mp + zx_pp) + 0.5 * (zm1 - zp1);
    const double uxzz = 0.5 * (xm1 - xp1) + 0.25 * (zx_pp - zx_mm - zx_pm + zx_mp);
    const double uxyz = 0.125 * (ppp - mpp - pmp - ppm + mmp + mpm + pmm - mmm);

    /* reconstruction.c:110-116 */
    const double is4 = ux * ux + uy * uy + uz * uz +
                       k_13_3 * (uxx * uxx + uyy * uyy + uzz * uzz) +
                       k_7_6 * (uxy * uxy + uyz * uyz + uxz * uxz) +
                       0.2 * (ux * uxxx + uy * uyyy + uz * uzzz) +
                       4.7 * (uxxy * uxxy + uxxz * uxxz + uxyy * uxyy + uxzz * uxzz + uyyz * uyyz + uyzz * uyzz) +
                       39.06 * (uxxx * uxxx + uyyy * uyyy + uzzz * uzzz) +
                       k_61_48 * uxyz * uxyz;

    double w4 = g4 / pow(is4 + k_eps, 4.0);
    double total = w4;

    /* seven quadratic stencils: central, +x, -x, +y, -y, +z, -z (reconstruction.c:121-203) */
    const double cx = 0.5 * (xp1 - xm1), cy = 0.5 * (yp1 - ym1), cz = 0.5 * (zp1 - zm1);
    const double cxx = 0.5 * (xm1 - 2.0 * u0 + xp1);
    const double cyy = 0.5 * (ym1 - 2.0 * u0 + yp1);
    const double czz = 0.5 * (zm1 - 2.0 * u0 + zp1);
    const double cxy = 0.25 * (xy_pp - xy_pm - xy_mp + xy_mm);
    const double cyz = 0.25 * (yz_pp - yz_pm - yz_mp + yz_mm);
    const double cxz = 0.25 * (zx_pp - zx_pm - zx_mp + zx_mm);

    double d[7][NQ];
    for (int s = 0; s < 7; ++s) {
        d[s][DX] = cx;  d[s][DY] = cy;  d[s][DZ] = cz;
        d[s][DXX] = cxx; d[s][DYY] = cyy; d[s][DZZ] = czz;
        d[s][DXY] = cxy; d[s][DYZ] = cyz; d[s][DXZ] = cxz;
    }
    /* +x */
    d[1][DX] = -1.5 * u0 + 2.0 * xp1 - 0.5 * xp2;
    d[1][DXX] = 0.5 * (u0 - 2.0 * xp1 + xp2);
    d[1][DXY] = 0.5 * (xy_pp - xy_pm) - d[1][DY];
    d[1][DXZ] = 0.5 * (zx_pp - zx_mp) - d[1][DZ];
    /* -x */
    d[2][DX] = -2.0 * xm1 + 0.5 * xm2 + 1.5 * u0;
    d[2][DXX] = 0.5 * (xm2 - 2.0 * xm1 + u0);
    d[2][DXY] = 0.5 * (xy_mm - xy_mp) + d[2][DY];
    d[2][DXZ] = 0.5 * (zx_mm - zx_pm) + d[2][DZ];
    /* +y */
    d[3][DY] = -1.5 * u0 + 2.0 * yp1 - 0.5 * yp2;
    d[3][DYY] = 0.5 * (u0 - 2.0 * yp1 + yp2);
    d[3][DXY] = 0.5 * (xy_pp - xy_mp) - d[3][DX];
    d[3][DYZ] = 0.5 * (yz_pp - yz_pm) - d[3][DZ];
    /* -y */
    d[4][DY] = -2.0 * ym1 + 0.5 * ym2 + 1.5 * u0;
    d[4][DYY] = 0.5 * (ym2 - 2.0 * ym1 + u0);
    d[4][DXY] = 0.5 * (xy_mm - xy_pm) + d[4][DX];
    d[4][DYZ] = 0.5 * (yz_mm - yz_mp) + d[4][DZ];
    /* +z */
    d[5][DZ] = -1.5 * u0 + 2.0 * zp1 - 0.5 * zp2;
    d[5][DZZ] = 0.5 * (u0 - 2.0 * zp1 + zp2);
    d[5][DYZ] = 0.5 * (yz_pp - yz_mp) - d[5][DY];
    d[5][DXZ] = 0.5 * (zx_pp - zx_pm) - d[5][DX];
    /* -z */
    d[6][DZ] = -2.0 * zm1 + 0.5 * zm2 + 1.5 * u0;
    d[6][DZZ] = 0.5 * (zm2 - 2.0 * zm1 + u0);
    d[6][DYZ] = 0.5 * (yz_mm - yz_pm) + d[6][DY];
    d[6][DXZ] = 0.5 * (zx_mm - zx_mp) + d[6][DX];

    double w3[7];
    for (int s = 0; s < 7; ++s) {   /* reconstruction.c:205-213 */
        w3[s] = g3[s] / pow(smooth_q(d[s]) + k_eps, 4.0);
        total += w3[s];
    }
    w4 = w4 / total;                /* reconstruction.c:215-218 */
    for (int s = 0; s < 7; ++s) w3[s] = w3[s] / total;

    double lin[NQ], nl[NQ];
    for (int q = 0; q < NQ; ++q) { lin[q] = 0.0; nl[q] = 0.0; }
    for (int s = 0; s < 7; ++s)     /* reconstruction.c:223-234 */
        for (int q = 0; q < NQ; ++q) {
            lin[q] += g3[s] * d[s][q];
            nl[q] += w3[s] * d[s][q];
        }

    const double wt = w4 / g4;      /* reconstruction.c:236 */
    co[0] = u0;                     /* reconstruction.c:238-257 */
    co[1] = wt * (ux - lin[DX]) + nl[DX];
    co[2] = wt * (uy - lin[DY]) + nl[DY];
    co[3] = wt * (uz - lin[DZ]) + nl[DZ];
    co[4] = wt * (uxx - lin[DXX]) + nl[DXX];
    co[5] = wt * (uyy - lin[DYY]) + nl[DYY];
    co[6] = wt * (uzz - lin[DZZ]) + nl[DZZ];
    co[7] = wt * (uxy - lin[DXY]) + nl[DXY];
    co[8] = wt * (uyz - lin[DYZ]) + nl[DYZ];
    co[9] = wt * (uxz - lin[DXZ]) + nl[DXZ];
    co[10] = wt * uxxx;
    co[11] = wt * uyyy;
    co[12] = wt * uzzz;
    co[13] = wt * uxxy;
    co[14] = wt * uxyy;
    co[15] = wt * uyyz;
    co[16] = wt * uyzz;
    co[17] = wt * uxxz;
    co[18] = wt * uxzz;
    co[19] = wt * uxyz;

    if (diag) {
        diag[0] = wt;
        diag[1] = w4;
        for (int s = 0; s < 7; ++s) diag[2 + s] = w3[s];
        diag[9] = is4;
    }
    if (dstencil)   /* the 9 derivative estimates of each quadratic stencil */
        for (int s = 0; s < 7; ++s)
            for (int q = 0; q < NQ; ++q) dstencil[s * NQ + q] = d[s][q];
    (void)k_1_3;
}

double orc_poly3d(double x, double y, double z, const double co[20])
{
    /* reconstruction.c:12-21 */
    return co[0] + co[1] * x + co[2] * y + co[3] * z +
           co[4] * (x * x - k_1_12) + co[5] * (y * y - k_1_12) + co[6] * (z * z - k_1_12) +
           co[7] * x * y + co[8] * y * z + co[9] * z * x +
           co[10] * x * (x * x - k_3_20) + co[11] * y * (y * y - k_3_20) + co[12] * z * (z * z - k_3_20) +
           co[13] * (x * x - k_1_12) * y + co[14] * x * (y * y - k_1_12) +
           co[15] * (y * y - k_1_12) * z + co[16] * y * (z * z - k_1_12) +
           co[17] * (x * x - k_1_12) * z + co[18] * x * (z * z - k_1_12) +
           co[19] * x * y * z;
}

void orc_grad3d(const double co[20], double x, double y, double z, double h,
                double *gx, double *gy, double *gz)
{
    /* reconstruction.c:24-34 */
    *gx = (co[1] + 2.0 * co[4] * x + co[7] * y + co[9] * z + co[10] * (3.0 * x * x - k_3_20) +
           2.0 * co[13] * x * y + co[14] * (y * y - k_1_12) + 2.0 * co[17] * x * z +
           co[18] * (z * z - k_1_12) + co[19] * y * z) / h;
    *gy = (co[2] + 2.0 * co[5] * y + co[7] * x + co[8] * z + co[11] * (3.0 * y * y - k_3_20) +
           co[13] * (x * x - k_1_12) + 2.0 * co[14] * x * y + 2.0 * co[15] * y * z +
           co[16] * (z * z - k_1_12) + co[19] * x * z) / h;
    *gz = (co[3] + 2.0 * co[6] * z + co[8] * y + co[9] * x + co[12] * (3.0 * z * z - k_3_20) +
           co[15] * (y * y - k_1_12) + 2.0 * co[16] * y * z + co[17] * (x * x - k_1_12) +
           2.0 * co[18] * x * z + co[19] * x * y) / h;
}

/* ---------------------------------------------------------------------- */

void orc_weno2d_linear(const double *c, ptrdiff_t sx, ptrdiff_t sy, double co[10])
{
#define V(a, b) c[(a) * sx + (b) * sy]
    const double u0 = V(0, 0);
    const double xm2 = V(-2, 0), xm1 = V(-1, 0), xp1 = V(1, 0), xp2 = V(2, 0);
    const double ym2 = V(0, -2), ym1 = V(0, -1), yp1 = V(0, 1), yp2 = V(0, 2);
    const double pp = V(1, 1), pm = V(1, -1), mp = V(-1, 1), mm = V(-1, -1);
#undef V
    /* 2d/weno4/standing_wave/reconstruction.c:114-123 */
    co[0] = u0;
    co[1] = k_41_60 * (-xm1 + xp1) + k_11_120 * (xm2 - xp2);
    co[2] = k_41_60 * (-ym1 + yp1) + k_11_120 * (ym2 - yp2);
    co[3] = -u0 + 0.5 * (xm1 + xp1);
    co[4] = -u0 + 0.5 * (ym1 + yp1);
    co[5] = 0.25 * (mm - mp - pm + pp);
    co[6] = k_1_6 * (xm1 - xp1) + k_1_12 * (xp2 - xm2);
    co[7] = k_1_6 * (ym1 - yp1) + k_1_12 * (yp2 - ym2);
    co[8] = 0.25 * (-mm + mp - pm + pp) + 0.5 * (ym1 - yp1);
    co[9] = 0.5 * (xm1 - xp1) + 0.25 * (pp - mm - mp + pm);
}

static double pow4(double a) { double a2 = a * a; return a2 * a2; }

void orc_cweno2d(const double *c, ptrdiff_t sx, ptrdiff_t sy, double gamma_hi, double gamma_lo,
                 double co[10])
{
    orc_cweno2d_ext(c, sx, sy, gamma_hi, gamma_lo, co, NULL, NULL);
}

void orc_cweno2d_ext(const double *c, ptrdiff_t sx, ptrdiff_t sy, double gamma_hi, double gamma_lo,
                     double co[10], double *diag, double *dstencil)
{
#define V(a, b) c[(a) * sx + (b) * sy]
    const double u0 = V(0, 0);
    const double xm2 = V(-2, 0), xm1 = V(-1, 0), xp1 = V(1, 0), xp2 = V(2, 0);
    const double ym2 = V(0, -2), ym1 = V(0, -1), yp1 = V(0, 1), yp2 = V(0, 2);
    const double pp = V(1, 1), pm = V(1, -1), mp = V(-1, 1), mm = V(-1, -1);
#undef V
    /* 2d/weno4/bubble/reconstruction.c:130-140 */
    const double g4 = gamma_hi;
    double g3[5];
    g3[0] = (1.0 - gamma_hi) * gamma_lo;
    for (int s = 1; s < 5; ++s) g3[s] = 0.25 * (1 - gamma_hi) * (1 - gamma_lo);

    /* :144-152 */
    const double ux = k_41_60 * (-xm1 + xp1) + k_11_120 * (xm2 - xp2);
    const double uy = k_41_60 * (-ym1 + yp1) + k_11_120 * (ym2 - yp2);
    const double uxx = -u0 + 0.5 * (xm1 + xp1);
    const double uyy = -u0 + 0.5 * (ym1 + yp1);
    const double uxy = 0.25 * (mm - mp - pm + pp);
    const double uxxx = k_1_6 * (xm1 - xp1) + k_1_12 * (xp2 - xm2);
    const double uyyy = k_1_6 * (ym1 - yp1) + k_1_12 * (yp2 - ym2);
    const double uxxy = 0.25 * (-mm + mp - pm + pp) + 0.5 * (ym1 - yp1);
    const double uxyy = 0.5 * (xm1 - xp1) + 0.25 * (pp - mm - mp + pm);

    /* :154-156 */
    const double is4 = (ux + 0.1 * uxxx) * (ux + 0.1 * uxxx) + k_13_3 * (uxx * uxx + uyy * uyy) +
                       (uy + 0.1 * uyyy) * (uy + 0.1 * uyyy) + 39.05 * (uxxx * uxxx + uyyy * uyyy) +
                       k_7_6 * uxy * uxy + 4.7 * (uxxy * uxxy + uxyy * uxyy);
    double w4 = g4 / pow4(is4 + k_eps);
    double sum = w4;

    /* five quadratic stencils: central, -x, +x, -y, +y (:161-199) */
    double dx[5], dy[5], dxx[5], dyy[5], dxy[5];
    dx[0] = 0.5 * (-xm1 + xp1);
    dy[0] = 0.5 * (-ym1 + yp1);
    dxx[0] = -u0 + 0.5 * (xm1 + xp1);
    dyy[0] = -u0 + 0.5 * (ym1 + yp1);
    dxy[0] = 0.25 * (mm - mp - pm + pp);

    dx[1] = 1.5 * u0 - 2.0 * xm1 + 0.5 * xm2;
    dy[1] = 0.5 * (-ym1 + yp1);
    dxx[1] = 0.5 * u0 - xm1 + 0.5 * xm2;
    dyy[1] = -u0 + 0.5 * ym1 + 0.5 * yp1;
    dxy[1] = 0.5 * (mm - mp - ym1 + yp1);

    dx[2] = -1.5 * u0 + 2.0 * xp1 - 0.5 * xp2;
    dy[2] = 0.5 * (-ym1 + yp1);
    dxx[2] = 0.5 * u0 - xp1 + 0.5 * xp2;
    dyy[2] = -u0 + 0.5 * ym1 + 0.5 * yp1;
    dxy[2] = 0.5 * (-pm + pp + ym1 - yp1);

    dx[3] = 0.5 * (-xm1 + xp1);
    dy[3] = 1.5 * u0 - 2.0 * ym1 + 0.5 * ym2;
    dxx[3] = -u0 + 0.5 * xm1 + 0.5 * xp1;
    dyy[3] = 0.5 * u0 - ym1 + 0.5 * ym2;
    dxy[3] = 0.5 * (-xm1 + mm + xp1 - pm);

    dx[4] = 0.5 * (-xm1 + xp1);
    dy[4] = -1.5 * u0 + 2.0 * yp1 - 0.5 * yp2;
    dxx[4] = -u0 + 0.5 * xm1 + 0.5 * xp1;
    dyy[4] = 0.5 * u0 - yp1 + 0.5 * yp2;
    dxy[4] = 0.5 * (xm1 - mp - xp1 + pp);

    double w3[5];
    for (int s = 0; s < 5; ++s) {   /* :203-207 */
        double is = dx[s] * dx[s] + dy[s] * dy[s] + k_13_3 * (dxx[s] * dxx[s] + dyy[s] * dyy[s]) +
                    k_7_6 * dxy[s] * dxy[s];
        w3[s] = g3[s] / pow4(is + k_eps);
        sum += w3[s];
    }
    w4 = w4 / sum;
    for (int s = 0; s < 5; ++s) w3[s] = w3[s] / sum;
    const double wt = w4 / g4;

    /* :219-249, the five-term sums written out left to right */
#define COMB(R4, D) (wt * (R4 - g3[0] * D[0] - g3[1] * D[1] - g3[2] * D[2] - g3[3] * D[3] - g3[4] * D[4]) + \
                     w3[0] * D[0] + w3[1] * D[1] + w3[2] * D[2] + w3[3] * D[3] + w3[4] * D[4])
    co[0] = u0;
    co[1] = COMB(ux, dx);
    co[2] = COMB(uy, dy);
    co[3] = COMB(uxx, dxx);
    co[4] = COMB(uyy, dyy);
    co[5] = COMB(uxy, dxy);
#undef COMB
    co[6] = wt * uxxx;
    co[7] = wt * uyyy;
    co[8] = wt * uxxy;
    co[9] = wt * uxyy;

    if (diag) {
        diag[0] = wt;
        diag[1] = w4;
        for (int s = 0; s < 5; ++s) diag[2 + s] = w3[s];
        diag[7] = is4;
    }
    if (dstencil)   /* x, y, xx, yy, xy of each quadratic stencil */
        for (int s = 0; s < 5; ++s) {
            dstencil[5 * s + 0] = dx[s];
            dstencil[5 * s + 1] = dy[s];
            dstencil[5 * s + 2] = dxx[s];
            dstencil[5 * s + 3] = dyy[s];
            dstencil[5 * s + 4] = dxy[s];
        }
}
