/*
 * oracle.h -- FP64 CPU restatement of the acoustic_petsc numerics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the acoustic_petsc source tree).  The arithmetic follows the reference's
 * evaluation order where that is cheap, so that differences against the GPU
 * kernels stay at the rounding level.
 */
#ifndef ACX_ORACLE_H
#define ACX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- operators -------------------------------------------------------- */

enum { ORC_WAVE2 = 0, ORC_ACOUSTIC4 = 1 };      /* system switch            */
/* per-axis boundary kind: ADIABATIC = odd reflection of component 0 (WAVE2,
 * 3d/weno4/plane_wave/rhs_function.c:48-135); REFLECTIVE / TRANSMISSIVE =
 * the LinearisedEuler walls (3d/weno3/LinearisedEuler/rhs_function.c:47-176,
 * ACOUSTIC4): mirror with the normal velocity negated / copy of the
 * boundary cell. */
enum { ORC_PERIODIC = 0, ORC_ADIABATIC = 1, ORC_REFLECTIVE = 2, ORC_TRANSMISSIVE = 3 };

typedef struct {
    int    N[3];        /* cells per axis (x, y, z)                          */
    double h;           /* cell size (uniform)                               */
    int    system;      /* ORC_WAVE2 (u,v) or ORC_ACOUSTIC4 (p,u,v,w)        */
    double gamma_hi;    /* CWENO linear weight of the 4th-order stencil      */
    double gamma_lo;    /* CWENO split of the rest (central vs sided)        */
    double c0;          /* wave speed                                        */
    double rho0;        /* base density (ACOUSTIC4)                          */
    double cfl;         /* CFL number used for the dt estimate (ACOUSTIC4)   */
    int    bc[3];       /* boundary kind per axis (both faces)               */
} orc_grid3d;

int orc_nvar(int system);

/* 3D CWENO4 reconstruction of one cell (3d/weno4/plane_wave/reconstruction.c:38-259).
 * `c` points at the cell's value inside a strided array; sx, sy, sz are the
 * element strides of one cell step along x, y, z.  The 33 values read are
 * the axis arms (+-1, +-2), the three planar diagonal sets and the 8 corners.
 * Writes the 20 Legendre-type coefficients (basis of evaluate_polynomial,
 * reconstruction.c:12-21).  diag (may be NULL) receives 10 numbers:
 * [0] wt_ratio, [1] w_R4, [2..8] w_R3[0..6], [9] IS_R4. */
void orc_cweno3d(const double *c, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                 double gamma_hi, double gamma_lo, double coeffs[20], double *diag);
/* The same, also writing (dstencil, may be NULL) the 9 derivative estimates
 * x, y, z, xx, yy, zz, xy, yz, xz of each of the 7 quadratic stencils in the
 * order central, +x, -x, +y, -y, +z, -z (reconstruction.c:121-203): 63 numbers. */
void orc_cweno3d_ext(const double *c, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                     double gamma_hi, double gamma_lo, double coeffs[20], double *diag, double *dstencil);

/* reconstruction.c:12-21 and :24-34 */
double orc_poly3d(double x, double y, double z, const double coeffs[20]);
void   orc_grad3d(const double coeffs[20], double x, double y, double z, double h,
                  double *gx, double *gy, double *gz);

/* Right-hand side with the reference's loop structure: ghosted local copy,
 * per-cell reconstruction into scratch at the 2x2 face Gauss points, per
 * Gauss point Riemann solves, then the flux divergence.
 *   WAVE2:     3d/weno4/plane_wave/rhs_function.c:12-345 (central flux pde.c:12-34)
 *   ACOUSTIC4: 3d/weno3/LinearisedEuler/rhs_function.c:178-384 with the CWENO4
 *              reconstruction above and PDEConsFlux/PDELLFRiemannSolver
 *              (3d/weno3/LinearisedEuler/pde.c:35-95).
 * U and RHS are DMDA global vectors: N[0]*N[1]*N[2] cells, interlaced dof,
 * x fastest.  dt_out (may be NULL) receives CFL*h/(3*s_max) (ACOUSTIC4,
 * rhs_function.c:382).  Returns 0, or 1 if a NaN/Inf state reaches a flux. */
int orc_rhs3d_faithful(const orc_grid3d *g, const double *U, double *RHS, double *dt_out);

/* Same operator, streaming form: each cell's face averages are computed once
 * from the reconstruction (the 2x2 Gauss rule collapses onto c0,c1,c4,c10 per
 * axis), fluxes of face averages, OpenMP over z.  Equal to the faithful form
 * up to rounding; used for large grids and for CPU timing. */
int orc_rhs3d_stream(const orc_grid3d *g, const double *U, double *RHS, int nthreads);

/* ---- PETSc TSSSP steppers (third-party: PETSc src/ts/impls/explicit/ssp/ssp.c) ---- */

typedef int (*orc_rhs_fn)(void *ctx, double t, const double *U, double *F);

/* TSSSPRKS3 with s = n^2 stages (Ketcheson 2008, low storage), one step. */
int orc_ssp_rks3_step(orc_rhs_fn f, void *ctx, size_t n, double *u, double t0, double dt,
                      int stages, double *w0, double *w1, double *F);
/* TSSSPRK104 (Ketcheson's SSPRK(10,4)), one step. */
int orc_ssp_rk104_step(orc_rhs_fn f, void *ctx, size_t n, double *u, double t0, double dt,
                       double *w0, double *w1, double *F);
/* TSSolve with TS_EXACTFINALTIME_MATCHSTEP: steps of dt, the last one cut to
 * land on t_final.  scheme 0 = RKS3 (stages), 1 = RK104.  Returns steps taken
 * (negative on RHS failure). */
int orc_ts_solve(orc_rhs_fn f, void *ctx, size_t n, double *u, double t0, double dt,
                 double t_final, int scheme, int stages);

/* RHS adaptor for orc_ts_solve: ctx is an orc_grid3d*; faithful / streaming. */
int orc_rhs3d_cb_faithful(void *ctx, double t, const double *U, double *F);

/* WENO point source and receiver (3d/weno4/point_source/rhs_function.c:337-346,
 * source.c:11-17): the faithful RHS plus Source(t)/h^3 on component 1 of cell
 * src; every call appends (t, component 0 of U at cell rcv) to the trace. */
typedef struct {
    orc_grid3d g;
    int    src[3], rcv[3];
    double f0;
    double *trace_t, *trace_u;   /* caller-owned, capacity cap */
    int    ntrace, cap;
} orc_point_source3d;
double orc_weno_source(double t, double f0);
int orc_rhs3d_cb_point_source(void *ctx, double t, const double *U, double *F);
int orc_rhs3d_cb_stream(void *ctx, double t, const double *U, double *F);

/* ---- 3D plane-wave convergence case (3d/weno4/plane_wave) --------------- */

/* InitializeSolution (initialize_solution.c:12-76) with its corner-offset
 * quadrature xGP = xc + l*h*x_gp5[l] (kept as written), WAVE2 plane wave
 * k=(1,1,1), c=1 (initial_condition.c:11-42), cell i centre at (i+1/2)h. */
void orc_plane_wave_init(int N, double h, double t, double *U);
/* ErrorNorms (error_norms.c:12-99): zeroes comp 1 of U in place, L2*h^1.5, Linf. */
void orc_plane_wave_error(int N, double h, double t, double *U, double *l2, double *linf);

/* ---- 2D siblings -------------------------------------------------------- */

/* 2d/weno4/standing_wave/reconstruction.c:106-125 (linear, 10 coeffs) */
void orc_weno2d_linear(const double *c, ptrdiff_t sx, ptrdiff_t sy, double coeffs[10]);
/* 2d/weno4/bubble/reconstruction.c:128-251 (CWENO4, gamma_hi/gamma_lo) */
void orc_cweno2d(const double *c, ptrdiff_t sx, ptrdiff_t sy, double gamma_hi, double gamma_lo,
                 double coeffs[10]);
/* The same with diag (may be NULL: [0] wt_ratio, [1] w_R4, [2..6] w_R3 of the
 * stencils central, -x, +x, -y, +y, [7] IS_R4) and dstencil (may be NULL:
 * x, y, xx, yy, xy of each of the 5 stencils in that order, 25 numbers). */
void orc_cweno2d_ext(const double *c, ptrdiff_t sx, ptrdiff_t sy, double gamma_hi, double gamma_lo,
                     double coeffs[10], double *diag, double *dstencil);

typedef struct {
    int    N[2];
    double h;
    int    system;      /* ORC_WAVE2 (u,v; standing_wave) or ORC_ACOUSTIC4 -> (p,u,v) */
    int    nonlinear;   /* 0: linear weno (standing_wave), 1: CWENO4 (bubble) */
    double gamma_hi, gamma_lo;
    double c0, rho0;
    int    bc[2];       /* ORC_PERIODIC / ORC_ADIABATIC (adiabatic_wall)      */
} orc_grid2d;

/* 2d/weno4/standing_wave/rhs_function.c:12-257 for WAVE2; for the 3-variable
 * acoustic system the LLF face flux of the 3D case restricted to 2D. */
int orc_rhs2d_faithful(const orc_grid2d *g, const double *U, double *RHS);
int orc_rhs2d_cb(void *ctx, double t, const double *U, double *F);

/* 2d/weno4/standing_wave: InitializeSolution (initialize_solution.c:12-53,
 * centred 5x5 Gauss), ErrorNorms (post_process.c:69-150, L2/N_x, Linf). */
void orc_standing_wave2d_init(int N, double *U);
void orc_standing_wave2d_error(int N, double t, double *U, double *l2, double *linf);

/* ---- CTCS leapfrog sibling (3d/ctcs/point_source/wave.c:138-240) ------- */

typedef struct {
    int    n[3];
    double dx, dy, dz, dt, c0;
    int    stencil;        /* 1: 2nd order, 2: 4th order                    */
    int    src[3], rcv[3]; /* source / receiver cell (-1: none)             */
    double f0;             /* Ricker-derivative dominant frequency          */
} orc_ctcs3d;

double orc_ctcs_source(double t, double f0);
/* One UpdatePressure call (wave.c:138-215); receiver value p(it) returned. */
double orc_ctcs3d_update(const orc_ctcs3d *c, const double *Pold, const double *P,
                         double *Pnew, const double *C, int it);
/* Runs nt steps from rest; trace[nt] gets the receiver samples. */
void orc_ctcs3d_run(const orc_ctcs3d *c, int nt, double *trace);

/* OpenMP thread count used by the oracle (0: library default). */
void orc_set_threads(int n);

/* ---- reductions ---------------------------------------------------------- */
double orc_norm2(size_t n, const double *a);
double orc_norminf(size_t n, const double *a);
/* relative L2 of a-b against b */
double orc_rel_l2(size_t n, const double *a, const double *b);

#ifdef __cplusplus
}
#endif
#endif
