/*
 * cases.c -- oracle restatement of the reference's convergence-study set-up
 * and error norms, and of the CTCS leapfrog sibling (test infrastructure
 * only; see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const double k_pi = 3.14159265358979323846264;   /* PETSC_PI */

/* five-point Gauss rule on [-1/2,1/2], hype.h:67-69 */
static const double gp5_x[5] = {0.0000000000000000, -0.2692346550528416, 0.2692346550528416,
                                -0.4530899229693320, 0.4530899229693320};
static const double gp5_w[5] = {0.28444444444444444, 0.23931433524968326, 0.23931433524968326,
                                0.11846344252809456, 0.11846344252809456};

/* DMDASetUniformCoordinates on a periodic axis: xmin + hx*i, hx = (xmax-xmin)/M;
 * on a non-periodic axis hx = (xmax-xmin)/(M-1). */
static double coord(double xmin, double xmax, int M, int periodic, int i)
{
    double hx = periodic ? (xmax - xmin) / M : (xmax - xmin) / (M - 1);
    return xmin + hx * i;
}

/* 3d/weno4/plane_wave/initial_condition.c:11-42, k=(1,1,1), wave_speed 1 */
static void plane_wave(double x, double y, double z, double t, double q[2])
{
    const double kx = 1.0, ky = 1.0, kz = 1.0;
    const double omega = 1.0 * sqrt(kx * kx + ky * ky + kz * kz);
    q[0] = sin(kx * x + ky * y + kz * z - omega * t);
    q[1] = -omega * cos(kx * x + ky * y + kz * z - omega * t);
}

/* initialize_solution.c:33-68 / error_norms.c:36-80: the quadrature point
 * offset is l*h*x_gp5[l] from the cell corner, exactly as written there. */
static void plane_wave_cell(int N, double h, int i, int j, int k, double t, double out[2])
{
    const double x0 = 0.0 + 0.5 * h, x1 = 2.0 * k_pi + 0.5 * h;   /* hype.c:115-117 */
    const double xc = coord(x0, x1, N, 1, i) - 0.5 * h;
    const double yc = coord(x0, x1, N, 1, j) - 0.5 * h;
    const double zc = coord(x0, x1, N, 1, k) - 0.5 * h;
    double acc[2] = {0.0, 0.0}, q[2];
    for (int l = 0; l < 5; l++)
        for (int m = 0; m < 5; m++)
            for (int n = 0; n < 5; ++n) {
                double xg = xc + (double)l * h * gp5_x[l];
                double yg = yc + (double)m * h * gp5_x[m];
                double zg = zc + (double)n * h * gp5_x[n];
                plane_wave(xg, yg, zg, t, q);
                for (int c = 0; c < 2; ++c) acc[c] += gp5_w[l] * gp5_w[m] * gp5_w[n] * q[c];
            }
    out[0] = acc[0];
    out[1] = acc[1];
}

void orc_plane_wave_init(int N, double h, double t, double *U)
{
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i)
                plane_wave_cell(N, h, i, j, k, t, U + (((size_t)k * N + j) * N + i) * 2);
}

void orc_plane_wave_error(int N, double h, double t, double *U, double *l2, double *linf)
{
    const size_t n = (size_t)N * N * N * 2;
    double *E = (double *)malloc(n * sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                size_t o = (((size_t)k * N + j) * N + i) * 2;
                double ex[2];
                plane_wave_cell(N, h, i, j, k, t, ex);
                E[o] = ex[0];
                E[o + 1] = 0.0;
                U[o + 1] = 0.0;   /* error_norms.c:72-74 zeroes it in place */
            }
    for (size_t e = 0; e < n; ++e) E[e] = E[e] + -1.0 * U[e];   /* VecAXPY(U_exact,-1,U) */
    *linf = orc_norminf(n, E);
    *l2 = orc_norm2(n, E) * pow(h, 1.5);
    free(E);
}

/* ---- 2d/weno4/standing_wave ------------------------------------------ */

static void standing_wave(double x, double y, double t, double q[2])
{
    /* 2d/weno4/standing_wave/initial_condition.c:11-35 */
    const double kx = 2 * k_pi, ky = 2 * k_pi;
    const double omega = 1.0 * sqrt(kx * kx + ky * ky);
    q[0] = sin(kx * x) * sin(ky * y) * cos(omega * t);
    q[1] = -omega * sin(kx * x) * sin(ky * y) * sin(omega * t);
}

static void standing_cell(int N, int i, int j, double t, double out[2])
{
    const double h = (1.0 - -1.0) / (double)N;   /* hype.c:49 */
    const double x0 = -1.0 + 0.5 * h, x1 = 1.0 - 0.5 * h;   /* hype.c:117-120 */
    const double xc = coord(x0, x1, N, 0, i), yc = coord(x0, x1, N, 0, j);
    double acc[2] = {0.0, 0.0}, q[2];
    for (int l = 0; l < 5; ++l)
        for (int m = 0; m < 5; ++m) {
            double xg = xc + h * gp5_x[l], yg = yc + h * gp5_x[m];
            standing_wave(xg, yg, t, q);
            for (int c = 0; c < 2; ++c) acc[c] += gp5_w[l] * gp5_w[m] * q[c];
        }
    out[0] = acc[0];
    out[1] = acc[1];
}

void orc_standing_wave2d_init(int N, double *U)
{
    /* initialize_solution.c:26-53; at t=0 the exact solution is the initial condition */
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
            standing_cell(N, i, j, 0.0, U + ((size_t)j * N + i) * 2);
}

void orc_standing_wave2d_error(int N, double t, double *U, double *l2, double *linf)
{
    /* post_process.c:69-150 */
    const size_t n = (size_t)N * N * 2;
    double *E = (double *)malloc(n * sizeof(double));
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            size_t o = ((size_t)j * N + i) * 2;
            double ex[2];
            standing_cell(N, i, j, t, ex);
            E[o] = ex[0];
            E[o + 1] = 0.0;
            U[o + 1] = 0.0;
        }
    for (size_t e = 0; e < n; ++e) E[e] = E[e] + -1.0 * U[e];
    *linf = orc_norminf(n, E);
    *l2 = orc_norm2(n, E) / ((double)N);
    free(E);
}

/* ---- 3d/ctcs/point_source/wave.c ------------------------------------- */

double orc_ctcs_source(double t, double f0)
{
    /* wave.c:234-240 */
    double t0 = 4.0 / f0;
    return -2.0 * (t - t0) * f0 * f0 * exp(-1.0 * f0 * f0 * (t - t0) * (t - t0));
}

double orc_ctcs3d_update(const orc_ctcs3d *c, const double *Pold, const double *P,
                         double *Pnew, const double *C, int it)
{
    const int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    const double invdx2 = 1.0 / (c->dx * c->dx), invdy2 = 1.0 / (c->dy * c->dy), invdz2 = 1.0 / (c->dz * c->dz);
    const double dt = c->dt;
    double rcv = 0.0;
    /* periodic DMDA (wave.c:86): ghost indices wrap */
#define W(a, n) (((a) + (n)) % (n))
#define PP(i, j, k) P[((size_t)W(k, nz) * ny + W(j, ny)) * nx + W(i, nx)]
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double pxx, pyy, pzz;
                const double p0 = PP(i, j, k);
                if (c->stencil == 1) {   /* wave.c:173-177 */
                    pxx = (PP(i + 1, j, k) - 2.0 * p0 + PP(i - 1, j, k)) * invdx2;
                    pyy = (PP(i, j + 1, k) - 2.0 * p0 + PP(i, j - 1, k)) * invdy2;
                    pzz = (PP(i, j, k + 1) - 2.0 * p0 + PP(i, j, k - 1)) * invdz2;
                } else {                 /* wave.c:178-183 */
                    pxx = (-1.0 / 12 * PP(i + 2, j, k) + 4.0 / 3.0 * PP(i + 1, j, k) - 5.0 / 2.0 * p0 +
                           4.0 / 3.0 * PP(i - 1, j, k) - 1.0 / 12.0 * PP(i - 2, j, k)) * invdx2;
                    pyy = (-1.0 / 12 * PP(i, j + 2, k) + 4.0 / 3.0 * PP(i, j + 1, k) - 5.0 / 2.0 * p0 +
                           4.0 / 3.0 * PP(i, j - 1, k) - 1.0 / 12.0 * PP(i, j - 2, k)) * invdy2;
                    pzz = (-1.0 / 12 * PP(i, j, k + 2) + 4.0 / 3.0 * PP(i, j, k + 1) - 5.0 / 2.0 * p0 +
                           4.0 / 3.0 * PP(i, j, k - 1) - 1.0 / 12.0 * PP(i, j, k - 2)) * invdz2;
                }
                const size_t o = ((size_t)k * ny + j) * nx + i;
                const double cc = C ? C[o] : c->c0;
                Pnew[o] = 2.0 * p0 - Pold[o] + cc * cc * dt * dt * (pxx + pyy + pzz);   /* wave.c:190 */
                if (i == c->src[0] && j == c->src[1] && k == c->src[2])                 /* wave.c:193-194 */
                    Pnew[o] += orc_ctcs_source(it * dt, c->f0) / (c->dx * c->dy * c->dz) * dt * dt;
                if (i == c->rcv[0] && j == c->rcv[1] && k == c->rcv[2]) rcv = p0;       /* wave.c:197-201 */
            }
#undef PP
#undef W
    return rcv;
}

void orc_ctcs3d_run(const orc_ctcs3d *c, int nt, double *trace)
{
    const size_t n = (size_t)c->n[0] * c->n[1] * c->n[2];
    double *Pold = (double *)calloc(n, sizeof(double));
    double *P = (double *)calloc(n, sizeof(double));
    double *Pnew = (double *)calloc(n, sizeof(double));
    for (int it = 0; it < nt; ++it) {   /* wave.c:109-121 */
        double r = orc_ctcs3d_update(c, Pold, P, Pnew, NULL, it);
        if (trace) trace[it] = r;
        memcpy(Pold, P, n * sizeof(double));
        memcpy(P, Pnew, n * sizeof(double));
    }
    free(Pold); free(P); free(Pnew);
}
