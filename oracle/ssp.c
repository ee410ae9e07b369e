/*
 * ssp.c -- oracle restatement of the strong-stability-preserving Runge-Kutta
 * steppers the reference selects (test infrastructure only; see oracle.h).
 *
 * The reference does not implement these itself: it calls PETSc's TSSSP
 * (3d/weno4/plane_wave/hype.c:180-182: TSSSPRKS3 with 4 stages;
 * 2d/weno4/standing_wave/hype.c:213-214: TSSSPRK104).  PETSc is neither
 * vendored nor installed, so the algorithms are restated from their
 * publication (D. Ketcheson, "Highly efficient strong stability preserving
 * Runge-Kutta methods with low-storage implementations", SISC 30(4), 2008:
 * the low-storage SSPRK(s,3) with s = n^2 and SSPRK(10,4)) in the vector
 * operation sequence PETSc's ssp.c uses (VecAXPY, VecAXPBYPCZ, VecAXPBY).
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static void axpy(size_t n, double *y, double a, const double *x)
{
    for (size_t i = 0; i < n; ++i) y[i] += a * x[i];
}

/* z = alpha x + beta y + gamma z */
static void axpbypcz(size_t n, double *z, double alpha, double beta, double gamma,
                     const double *x, const double *y)
{
    for (size_t i = 0; i < n; ++i) z[i] = alpha * x[i] + beta * y[i] + gamma * z[i];
}

static double stage_frac(int i, int n, int s)
{
    /* stage abscissa of SSPRK(n^2,3) stage i */
    return (i < n * (n + 1) / 2) ? 1. * i / (s - n) : (1. * i - n) / (s - n);
}

int orc_ssp_rks3_step(orc_rhs_fn f, void *ctx, size_t n_el, double *u, double t0, double dt,
                      int s, double *w0, double *w1, double *F)
{
    const int n = (int)(sqrt((double)s) + 0.001);
    const int r = s - n;
    if (n * n != s) return -1;
    int i = 0, rc;
    memcpy(w0, u, n_el * sizeof(double));
    for (; i < (n - 1) * (n - 2) / 2; ++i) {
        if ((rc = f(ctx, t0 + stage_frac(i, n, s) * dt, w0, F))) return rc;
        axpy(n_el, w0, dt / r, F);
    }
    memcpy(w1, w0, n_el * sizeof(double));
    for (; i < n * (n + 1) / 2 - 1; ++i) {
        if ((rc = f(ctx, t0 + stage_frac(i, n, s) * dt, w0, F))) return rc;
        axpy(n_el, w0, dt / r, F);
    }
    if ((rc = f(ctx, t0 + stage_frac(i, n, s) * dt, w0, F))) return rc;
    axpbypcz(n_el, w0, 1. * n / (2 * n - 1), (n - 1.) * dt / (r * (2 * n - 1)), (n - 1.) * 1. / (2 * n - 1), w1, F);
    ++i;
    for (; i < s; ++i) {
        if ((rc = f(ctx, t0 + stage_frac(i, n, s) * dt, w0, F))) return rc;
        axpy(n_el, w0, dt / r, F);
    }
    memcpy(u, w0, n_el * sizeof(double));
    return 0;
}

int orc_ssp_rk104_step(orc_rhs_fn f, void *ctx, size_t n_el, double *u, double t0, double dt,
                       double *w0, double *w1, double *F)
{
    static const double c[10] = {0, 1. / 6, 2. / 6, 3. / 6, 4. / 6, 2. / 6, 3. / 6, 4. / 6, 5. / 6, 1};
    int i, rc;
    memcpy(w0, u, n_el * sizeof(double));
    for (i = 0; i < 5; ++i) {
        if ((rc = f(ctx, t0 + c[i] * dt, w0, F))) return rc;
        axpy(n_el, w0, dt / 6, F);
    }
    for (size_t e = 0; e < n_el; ++e) w1[e] = 1. / 25 * u[e] + 9. / 25 * w0[e];
    for (size_t e = 0; e < n_el; ++e) w0[e] = 15 * w1[e] + -5 * w0[e];
    for (; i < 9; ++i) {
        if ((rc = f(ctx, t0 + c[i] * dt, w0, F))) return rc;
        axpy(n_el, w0, dt / 6, F);
    }
    if ((rc = f(ctx, t0 + dt, w0, F))) return rc;
    axpbypcz(n_el, w1, 3. / 5, dt / 10, 1, w0, F);
    memcpy(u, w1, n_el * sizeof(double));
    return 0;
}

int orc_ts_solve(orc_rhs_fn f, void *ctx, size_t n, double *u, double t0, double dt,
                 double t_final, int scheme, int stages)
{
    double *w0 = (double *)malloc(n * sizeof(double));
    double *w1 = (double *)malloc(n * sizeof(double));
    double *F = (double *)malloc(n * sizeof(double));
    if (!w0 || !w1 || !F) { free(w0); free(w1); free(F); return -2; }
    double t = t0, h = dt;
    int steps = 0;
    const double small = 1.0e-10;   /* PETSC_SMALL for double */
    if (t + h > t_final) h = t_final - t;
    while (t < t_final) {
        int rc = scheme == 0 ? orc_ssp_rks3_step(f, ctx, n, u, t, h, stages, w0, w1, F)
                             : orc_ssp_rk104_step(f, ctx, n, u, t, h, w0, w1, F);
        if (rc) { steps = -1; break; }
        t += h;
        ++steps;
        /* TS_EXACTFINALTIME_MATCHSTEP: shorten the next step to land on t_final */
        if (t + h > t_final) {
            double next = t_final - t;
            if (next > small) h = next;
            else break;
        }
    }
    free(w0); free(w1); free(F);
    return steps;
}

double orc_norm2(size_t n, const double *a)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * a[i];
    return sqrt(s);
}

double orc_norminf(size_t n, const double *a)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) if (fabs(a[i]) > m) m = fabs(a[i]);
    return m;
}

double orc_rel_l2(size_t n, const double *a, const double *b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double d = a[i] - b[i];
        num += d * d;
        den += b[i] * b[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}
