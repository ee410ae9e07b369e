/*
 * rhs2d.c -- oracle restatement of the 2D right-hand sides (test
 * infrastructure only; see oracle.h).
 *
 *   WAVE2 (u, v):  2d/weno4/standing_wave/rhs_function.c:12-257 with the basis
 *                  tables of hype.c:146-170 and the central flux pde.c:12-34.
 *   acoustic (p, u, v): the 3D LinearisedEuler LLF face flux
 *                  (3d/weno3/LinearisedEuler/pde.c:35-95) on the 2D skeleton,
 *                  2 Gauss points per face (w = 1/2), reconstruction by the 2D
 *                  CWENO4 of 2d/weno4/bubble/reconstruction.c:128-251.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define GW 3

static const double gp2_x[2] = {-0.28867513459481287, 0.28867513459481287};   /* hype.h:55-57 */
static const double gp2_w[2] = {0.50000000000000000, 0.50000000000000000};

/* 2d/weno4/standing_wave/reconstruction.c:11-100 */
static double basis2d(double x, double y, int n)
{
    switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return y;
    case 3: return x * x - 1. / 12.;
    case 4: return y * y - 1. / 12.;
    case 5: return x * y;
    case 6: return x * (x * x - 3. / 20.);
    case 7: return y * (y * y - 3. / 20.);
    case 8: return y * (x * x - 1. / 12.);
    case 9: return x * (y * y - 1. / 12.);
    default: return 0.0;
    }
}

static void basis2d_grad(double x, double y, int n, double *gx, double *gy)
{
    switch (n) {
    case 0: *gx = 0.0; *gy = 0.0; break;
    case 1: *gx = 1.0; *gy = 0.0; break;
    case 2: *gx = 0.0; *gy = 1.0; break;
    case 3: *gx = 2.0 * x; *gy = 0.0; break;
    case 4: *gx = 0.0; *gy = 2.0 * y; break;
    case 5: *gx = y; *gy = x; break;
    case 6: *gx = 3.0 * x * x - 3. / 20.; *gy = 0.0; break;
    case 7: *gx = 0.0; *gy = 3.0 * y * y - 3. / 20.; break;
    case 8: *gx = 2.0 * x * y; *gy = (x * x - 1. / 12.); break;
    case 9: *gx = (y * y - 1. / 12.); *gy = 2.0 * x * y; break;
    default: *gx = 0.0; *gy = 0.0;
    }
}

/* face f (0 -x, 1 +x, 2 -y, 3 +y), Gauss point q */
static void face_pt(int f, int q, double *x, double *y)
{
    switch (f) {
    case 0: *x = -0.5; *y = gp2_x[q]; break;
    case 1: *x = 0.5; *y = gp2_x[q]; break;
    case 2: *x = gp2_x[q]; *y = -0.5; break;
    default: *x = gp2_x[q]; *y = 0.5; break;
    }
}

int orc_rhs2d_faithful(const orc_grid2d *g, const double *U, double *RHS)
{
    const int wave = (g->system == ORC_WAVE2);
    const int nv = wave ? 2 : 3;
    const int Nx = g->N[0], Ny = g->N[1];
    const int Gx = Nx + 2 * GW, Gy = Ny + 2 * GW;
    const int Rx = Nx + 2, Ry = Ny + 2;
    const double r1_h = 1. / g->h;
    int bad = 0;

    double *L = (double *)calloc((size_t)Gx * Gy * nv, sizeof(double));
    const size_t per_cell = wave ? 4 * 2 * 2 : (size_t)nv * 4 * 2;
    double *S = (double *)malloc((size_t)Rx * Ry * per_cell * sizeof(double));
    double *Fx = (double *)malloc((size_t)Ny * (Nx + 1) * nv * sizeof(double));
    double *Fy = (double *)malloc((size_t)(Ny + 1) * Nx * nv * sizeof(double));
    if (!L || !S || !Fx || !Fy) { free(L); free(S); free(Fx); free(Fy); return 2; }

    /* basis tables at the face Gauss points (hype.c:146-170) */
    double phi[4][2][10], gphx[4][2][10], gphy[4][2][10];
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 10; ++k)
            for (int f = 0; f < 4; ++f) {
                double x, y;
                face_pt(f, q, &x, &y);
                phi[f][q][k] = basis2d(x, y, k);
                basis2d_grad(x, y, k, &gphx[f][q][k], &gphy[f][q][k]);
            }

#define LOC(i, j) (L + ((size_t)((j) + GW) * Gx + ((i) + GW)) * nv)
#define SCR(i, j) (S + ((size_t)((j) + 1) * Rx + ((i) + 1)) * per_cell)
    /* global -> local with periodic wrap on periodic axes */
    for (int jj = 0; jj < Gy; ++jj)
        for (int ii = 0; ii < Gx; ++ii) {
            int i = ii - GW, j = jj - GW;
            if (i < 0 || i >= Nx) { if (g->bc[0] != ORC_PERIODIC) continue; i = (i + Nx) % Nx; }
            if (j < 0 || j >= Ny) { if (g->bc[1] != ORC_PERIODIC) continue; j = (j + Ny) % Ny; }
            memcpy(L + ((size_t)jj * Gx + ii) * nv, U + ((size_t)j * Nx + i) * nv, nv * sizeof(double));
        }
    /* adiabatic walls, x then y (rhs_function.c:50-118) */
    if (g->bc[0] == ORC_ADIABATIC)
        for (int j = -GW; j < Ny + GW; ++j)
            for (int i = -GW; i < Nx + GW; ++i) {
                if (i < 0) LOC(i, j)[0] = -LOC(-1 - i, j)[0];
                if (i >= Nx) LOC(i, j)[0] = -LOC(2 * (Nx - 1) - i + 1, j)[0];
            }
    if (g->bc[1] == ORC_ADIABATIC)
        for (int j = -GW; j < Ny + GW; ++j)
            for (int i = -GW; i < Nx + GW; ++i) {
                if (j < 0) LOC(i, j)[0] = -LOC(i, -1 - j)[0];
                if (j >= Ny) LOC(i, j)[0] = -LOC(i, 2 * (Ny - 1) - j + 1)[0];
            }

    /* reconstruction (rhs_function.c:124-173) */
    for (int j = -1; j < Ny + 1; ++j)
        for (int i = -1; i < Nx + 1; ++i) {
            double co[10];
            double *s = SCR(i, j);
            for (int c = 0; c < (wave ? 1 : nv); ++c) {
                if (g->nonlinear) orc_cweno2d(LOC(i, j) + c, nv, (ptrdiff_t)Gx * nv, g->gamma_hi, g->gamma_lo, co);
                else orc_weno2d_linear(LOC(i, j) + c, nv, (ptrdiff_t)Gx * nv, co);
                for (int f = 0; f < 4; ++f)
                    for (int q = 0; q < 2; ++q) {
                        if (wave) {
                            double gx = 0.0, gy = 0.0;
                            for (int k = 0; k < 10; ++k) {
                                gx += co[k] * gphx[f][q][k];
                                gy += co[k] * gphy[f][q][k];
                            }
                            gx = r1_h * gx; gy = r1_h * gy;
                            s[(f * 2 + q) * 2 + 0] = gx;
                            s[(f * 2 + q) * 2 + 1] = gy;
                        } else {
                            double v = 0.0;
                            for (int k = 0; k < 10; ++k) v += co[k] * phi[f][q][k];
                            s[(c * 4 + f) * 2 + q] = v;
                        }
                    }
            }
        }

    /* fluxes (rhs_function.c:175-237) */
    const double K = g->rho0 * g->c0 * g->c0, ir = 1.0 / g->rho0;
    for (int dir = 0; dir < 2; ++dir) {
        const int ex = dir == 0, ey = dir == 1;
        const double nx = ex, ny = ey;
        const int fl = 2 * dir + 1, fr = 2 * dir;
        double *Fd = dir == 0 ? Fx : Fy;
        const int Mx = Nx + ex, My = Ny + ey;
        for (int j = 0; j < My; ++j)
            for (int i = 0; i < Mx; ++i) {
                const double *sl = SCR(i - ex, j - ey), *sr = SCR(i, j);
                double *out = Fd + ((size_t)j * Mx + i) * nv;
                if (wave) {
                    double flux = 0.0;
                    for (int q = 0; q < 2; ++q) {
                        const double *gl = sl + (fl * 2 + q) * 2, *gr = sr + (fr * 2 + q) * 2;
                        double FL = g->c0 * g->c0 * (gl[0] * nx + gl[1] * ny);
                        double FR = g->c0 * g->c0 * (gr[0] * nx + gr[1] * ny);
                        flux += gp2_w[q] * (0.5 * (FR + FL));
                    }
                    out[0] = flux;
                } else {
                    double acc[3] = {0.0, 0.0, 0.0};
                    for (int q = 0; q < 2; ++q) {
                        double QL[3], QR[3], FL[3], FR[3];
                        for (int c = 0; c < 3; ++c) {
                            QL[c] = sl[(c * 4 + fl) * 2 + q];
                            QR[c] = sr[(c * 4 + fr) * 2 + q];
                            if (!isfinite(QL[c]) || !isfinite(QR[c])) bad = 1;
                        }
                        FL[0] = nx * K * QL[1] + ny * K * QL[2];
                        FL[1] = ir * nx * QL[0];
                        FL[2] = ir * ny * QL[0];
                        FR[0] = nx * K * QR[1] + ny * K * QR[2];
                        FR[1] = ir * nx * QR[0];
                        FR[2] = ir * ny * QR[0];
                        const double s = g->c0;
                        for (int c = 0; c < 3; ++c)
                            acc[c] += gp2_w[q] * (0.5 * (FR[c] + FL[c] - s * (QR[c] - QL[c])));
                    }
                    for (int c = 0; c < 3; ++c) out[c] = acc[c];
                }
            }
    }

    /* divergence (rhs_function.c:241-251) */
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            double *r = RHS + ((size_t)j * Nx + i) * nv;
            const double *fx0 = Fx + ((size_t)j * (Nx + 1) + i) * nv, *fx1 = fx0 + nv;
            const double *fy0 = Fy + ((size_t)j * Nx + i) * nv, *fy1 = fy0 + (size_t)Nx * nv;
            if (wave) {
                r[0] = LOC(i, j)[1];
                r[1] = r1_h * (fx1[0] - fx0[0]) + r1_h * (fy1[0] - fy0[0]);
            } else {
                for (int c = 0; c < nv; ++c) r[c] = -r1_h * (fx1[c] - fx0[c]) - r1_h * (fy1[c] - fy0[c]);
            }
        }
#undef LOC
#undef SCR
    free(L); free(S); free(Fx); free(Fy);
    return bad;
}

int orc_rhs2d_cb(void *ctx, double t, const double *U, double *F)
{
    (void)t;
    return orc_rhs2d_faithful((const orc_grid2d *)ctx, U, F);
}
