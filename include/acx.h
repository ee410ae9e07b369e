/*
 * acx.h -- C ABI of the MI355X (gfx950) kernels behind the acoustic_petsc
 * right-hand side and SSP stepper.
 *
 * Plain C: device pointers, sizes and a hipStream_t passed as void*.  No
 * torch or C++ types cross this boundary.  Every state array is a DMDA
 * global vector: N[0]*N[1]*N[2] cells, the components of a cell interlaced,
 * x fastest (the layout of DMCreateGlobalVector on one rank,
 * 3d/weno4/plane_wave/hype.c:102-111).
 *
 * Return codes: 0 ok, ACX_EINVAL bad argument, ACX_EUNSUPPORTED combination
 * not implemented, ACX_EHIP runtime failure, ACX_ENONFINITE a NaN/Inf state
 * was detected.  acx_last_error() describes the last failure of the calling
 * thread.
 */
#ifndef ACX_H
#define ACX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACX_OK             0
#define ACX_EINVAL         1
#define ACX_EUNSUPPORTED   2
#define ACX_EHIP           3
#define ACX_ENONFINITE     4

/* operator families (the reference's nVar 2 / nVar 4 programs) */
#define ACX_WAVE2      0   /* (u, v): u_t = v, v_t = div(c^2 grad u); 3d/weno4/plane_wave */
#define ACX_ACOUSTIC4  1   /* (p, u, v, w) linearised acoustics, LLF flux; 3d/weno3/LinearisedEuler */

#define ACX_PERIODIC      0   /* DM_BOUNDARY_PERIODIC */
#define ACX_ADIABATIC     1   /* WAVE2: DM_BOUNDARY_GHOSTED + odd reflection of component 0
                                 (3d/weno4/plane_wave/rhs_function.c:48-135) */
#define ACX_REFLECTIVE    2   /* ACOUSTIC4: mirror, normal velocity negated
                                 (3d/weno3/LinearisedEuler/rhs_function.c:47-176) */
#define ACX_TRANSMISSIVE  3   /* ACOUSTIC4: copy of the boundary cell (same file) */

/* 2D grids: n[2] = 1 with bc[2] = ACX_PERIODIC is one plane, the DMDA of the
 * reference's 2D programs.  The operator on it is the 2D WENO4 operator
 * (2d/weno4/standing_wave/rhs_function.c:12-257 for ACX_WAVE2; for
 * ACX_ACOUSTIC4 with w = 0 the 2D linearised acoustics (p', u, v), w staying
 * 0); with other than the shipped linear weights the reconstruction is the 2D
 * CWENO4 of 2d/weno4/bubble/reconstruction.c:128-251 (five quadratic
 * stencils). */
typedef struct acx_grid {
    int32_t n[3];        /* cells along x, y, z (every axis >= 4, or n[2] = 1) */
    int32_t system;      /* ACX_WAVE2 or ACX_ACOUSTIC4                         */
    int32_t bc[3];       /* boundary kind per axis (both faces of the axis)   */
    int32_t pad_;
    double  h;           /* uniform cell size                                  */
    double  gamma_hi;    /* CWENO linear weight of the 4th-order polynomial;
                            1.0 with gamma_lo = 0.0 is the shipped, linear
                            scheme (3d/weno4/plane_wave/reconstruction.c:42-43) */
    double  gamma_lo;
    double  c0;          /* wave speed                                         */
    double  rho0;        /* base density (ACOUSTIC4)                           */
} acx_grid;

/* Right-hand side F(U) of the semi-discrete operator, FP64.
 * Replaces RHSFunction (3d/weno4/plane_wave/rhs_function.c:12-345 for
 * ACX_WAVE2; 3d/weno3/LinearisedEuler/rhs_function.c:12-387 with the CWENO4
 * reconstruction of 3d/weno4/plane_wave/reconstruction.c:38-259 for
 * ACX_ACOUSTIC4).  U and RHS are device pointers of n[0]*n[1]*n[2]*nvar doubles. */
int acx_rhs_f64(const acx_grid *g, const double *U, double *RHS, void *stream);

/* One fused SSP stage: OUT = a*U0 + b*Q + c*F(Q) (FP64).  U0 may be NULL when
 * a == 0.  OUT must not alias Q or U0.  This is the VecAXPY /
 * VecAXPBYPCZ step of PETSc's TSSSP fused into the RHS evaluation
 * (PETSc ssp.c, called from 3d/weno4/plane_wave/hype.c:180-183). */
int acx_stage_f64(const acx_grid *g, const double *Q, const double *U0, double *OUT,
                  double a, double b, double c, void *stream);

/* One step of TSSSPRKS3 with 4 stages (hype.c:181-182): U <- step(U, dt).
 * W0, W1 are device work vectors of the same size as U; their contents after
 * the call are unspecified.  Four fused stage launches; on periodic one-plane
 * ACOUSTIC4 grids with the shipped weights and up to 2^20 cells, one launch
 * of the whole step into W0 and a copy back into U. */
int acx_ssp_rks3_step_f64(const acx_grid *g, double *U, double *W0, double *W1, double dt, void *stream);

/* nsteps steps of acx_ssp_rks3_step_f64 replayed from a hipGraph of their
 * stage launches (the reference's stage loop, hype.c:180-183, PETSc's
 * TSStep_SSP): a graph of min(nsteps, 32) steps is captured for a (device,
 * grid, U, W0, W1, dt) and replayed, the remainder from a graph of its own
 * size (one cached graph per device, captured again when any of these
 * change).  Results equal acx_ssp_rks3_step_f64 bit for bit. */
int acx_ssp_rks3_steps_graph_f64(const acx_grid *g, double *U, double *W0, double *W1, double dt, int nsteps,
                                 void *stream);

/* ---- FP32 storage with compensated step accumulation ------------------- *
 * The state is kept as a float pair (hi, lo) with hi + lo the value; stage
 * intermediates are FP32 increments D_k = q_k - (hi + lo).  One step of the
 * 4-stage SSP RK3 in that form: every stage reads hi and the current
 * increment, reconstructs in FP32, and the final update adds the step
 * increment to (hi, lo) with an error-free two-sum.  D0, D1 are device work
 * arrays of n*nvar floats. */
int acx_ssp_rks3_step_f32c(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                           float *LO_OUT, float *D0, float *D1, double dt, void *stream);

/* ---- reductions (FP64 accumulation) ---------------------------------- */

/* out[0] = sum of squares of component `comp` (all components if comp < 0),
 * out[1] = max abs; host pointer out[2].  VecNorm (error_norms.c:87-89).
 * A NaN anywhere makes both results NaN (as PETSc's VecNorm).  The partial
 * sums live in a buffer taken per call on `stream` (stream-ordered), so
 * calls on different streams do not share scratch. */
int acx_norms_f64(const double *X, size_t ncells, int nvar, int comp, double *out, void *stream);
/* Y = X for n doubles, stream-ordered (VecCopy, used by the SSP stages of
 * PETSc's TSSSP, which the reference selects at 3d/weno4/plane_wave/
 * hype.c:180-182): nontemporal 16-byte loads and stores, the state-sized
 * copy rate the stage kernels are compared with.  Overlapping distinct
 * arrays are refused (ACX_EINVAL); X == Y is a no-op. */
int acx_vec_copy_f64(const double *X, double *Y, size_t n, void *stream);
/* relative L2 of (A-B) against B and max |A-B|, FP64 accumulation; A is
 * float hi (+ lo, may be NULL), B double. */
int acx_diff_f32c_f64(const float *HI, const float *LO, const double *B, size_t n,
                      double *out, void *stream);

/* acoustic energy 1/2 sum (p^2/K + rho0 |u|^2) h^3 (ACOUSTIC4; h^2 on a 2D grid) */
int acx_energy_f64(const acx_grid *g, const double *U, double *out, void *stream);

/* CFL time step, 3d/weno3/LinearisedEuler/rhs_function.c:382-384:
 * CFL*h/(3*s_max), s_max = c0 for the acoustic LLF flux; CFL*h/(2*s_max) on
 * a 2D grid (2d/weno4/bubble/rhs_function.c:385). */
double acx_cfl_dt(const acx_grid *g, double cfl);

/* ---- CTCS leapfrog sibling, 3d/ctcs/point_source/wave.c:138-215 ------- */

typedef struct acx_ctcs {
    int32_t n[3];
    int32_t stencil;     /* 1: 2nd-order, 2: 4th-order Laplacian                */
    int32_t src[3];      /* source cell (-1 for none)                           */
    int32_t rcv[3];      /* receiver cell (-1 for none)                         */
    double  dx, dy, dz, dt;
    double  f0;          /* Ricker-derivative dominant frequency                */
} acx_ctcs;

/* One axis of 1 cell makes the 2D and 1D programs (2d/ctcs/plane_wave/wave.c:161-202,
 * 1d/ctcs/travelling_wave/wave.c:170-212): its second difference vanishes;
 * set its spacing to 1 so that the source divides by dx dy (or dx).
 * PNEW = 2P - POLD + c^2 dt^2 Lap(P) + source(it*dt)*dt^2/(dx dy dz) at src;
 * C is the per-cell wave speed (device); *rcv_out (host, may be NULL) gets
 * P at the receiver before the update. */
int acx_ctcs_update_f64(const acx_ctcs *c, const double *POLD, const double *P, double *PNEW,
                        const double *C, int it, double *rcv_out, void *stream);

/* ---- host-side helpers (no GPU needed) -------------------------------- */

/* Seeded synthetic fields on the host (splitmix64), exact cell averages:
 * a sum of `npulse` Gaussian pressure pulses on the periodic unit cube,
 * amplitudes U[amin,amax], sigma U[smin,smax], centres U[0,1)^3, zero
 * velocity.  U is n^3*nvar doubles (AoS). */
int acx_field_gaussians(const acx_grid *g, uint64_t seed, int npulse, double amin, double amax,
                        double smin, double smax, double *U);
/* Standing acoustic plane wave p = A sin(2 pi k.x) cos(wt), u = -(A/rho0 c0) khat cos(2pi k.x) sin(wt),
 * cell averages at time t (ACOUSTIC4, periodic unit cube). */
int acx_field_plane_wave(const acx_grid *g, const int k[3], double A, double t, double *U);

/* One entry of a device state: X[cell (i, j, k), component comp] += value.
 * With X the RHS and value = Source(t) / h^3 on component 1 this is the
 * point source of the reference's WENO driver
 * (3d/weno4/point_source/rhs_function.c:337-340). */
int acx_point_add_f64(const acx_grid *g, double *X, int i, int j, int k, int comp, double value, void *stream);

/* *out = X[cell (i, j, k), component comp] (host result; synchronises the
 * stream), the receiver read of 3d/weno4/point_source/rhs_function.c:342-346. */
int acx_point_get_f64(const acx_grid *g, const double *X, int i, int j, int k, int comp, double *out, void *stream);

/* ---- non-finite watch (ACX_ENONFINITE) -------------------------------- *
 * The reference aborts on an inadmissible state (3d/weno3/LinearisedEuler/
 * pde.c:50-55, MPI_Abort when p < 1e-12); here every stage kernel raises a
 * per-device flag when it writes a NaN or +-Inf.  The flag is read, without
 * synchronising, at the entry of every RHS / stage / step call (acx_rhs_f64,
 * acx_stage_f64, the steppers, the selective and ensemble steps), which then
 * return ACX_ENONFINITE without launching: a blow-up is reported by the first
 * such call issued after the kernel that produced it has finished.
 * acx_nonfinite_check synchronises `stream` first and so reports everything
 * queued on it before the call.  The flag stays raised (for every stream and
 * state on the device) until acx_nonfinite_clear. */
int acx_nonfinite_check(void *stream);
int acx_nonfinite_clear(void);

const char *acx_last_error(void);
int acx_version(void);

#ifdef __cplusplus
}
#endif
#endif
