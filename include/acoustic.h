/*
 * acoustic.h -- the drop-in right-hand side of the acoustic_petsc WENO-4
 * solvers, over petsclite.h.
 *
 * In the reference every program compiles its own RHSFunction
 * (3d/weno4/plane_wave/rhs_function.c:12-345; 3d/weno3/LinearisedEuler/
 * rhs_function.c:12-387) against a per-program hype.h that defines nVar, the
 * Field record and the AppCtx the driver fills (3d/weno4/plane_wave/hype.h:
 * 33-101, 236-266).  This header plays hype.h's role for the hot path: a
 * driver fills the same AppCtx fields, registers RHSFunction with
 * TSSetRHSFunction (hype.c:171) and calls TSSolve; the right-hand side and
 * the SSP stages then run as fused gfx950 kernels on device-resident vectors.
 */
#ifndef ACOUSTIC_H
#define ACOUSTIC_H

#include "petsclite.h"

#ifndef nVar
#define nVar 4               /* 2: (u, v) wave system, 4: (p, u, v, w) acoustics */
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    PetscReal comp[nVar];
} Field;

/* boundary kinds of the reference programs: periodic, adiabatic (WAVE2,
 * 3d/weno4/plane_wave/hype.h:115; 2D: adiabatic_wall), reflective and
 * transmissive (ACOUSTIC4, 3d/weno3/LinearisedEuler/hype.h:111).  Both faces
 * of a non-periodic axis take the same kind here. */
enum bndry_type { periodic, adiabatic, reflective, transmissive };

/* system selector: which of the reference's operators RHSFunction applies */
enum acx_system_type { WAVE_SYSTEM = 0, ACOUSTIC_SYSTEM = 1 };

typedef struct {
    /* fields of the reference AppCtx (3d/weno4/plane_wave/hype.h:236-266) */
    PetscReal x_min, y_min, z_min, x_max, y_max, z_max;
    PetscInt N_x, N_y, N_z;
    PetscReal CFL;
    PetscReal dt;              /* ACOUSTIC: set by RHSFunction to CFL h / (3 s_max),
                                  3d/weno3/LinearisedEuler/rhs_function.c:382-384 */
    PetscReal h;
    PetscReal InitialStep, InitialTime, FinalTime;
    PetscInt WriteInterval, RestartInterval;
    PetscBool Restart;
    enum bndry_type left_boundary, right_boundary, top_boundary, bottom_boundary, front_boundary, back_boundary;
    Vec localU;
    /* operator parameters that the reference hard-codes per program */
    enum acx_system_type system;   /* WAVE_SYSTEM: 3d/weno4, ACOUSTIC_SYSTEM: LinearisedEuler fluxes */
    PetscReal gammaHi, gammaLo;    /* reconstruction.c:42-43 (shipped 1, 0; nonlinear 0.85, 0.85) */
    PetscReal wave_speed;          /* c0 (hype.h:39 wave_speed, LinearisedEuler hype.h:37 c_0) */
    PetscReal rho_0;               /* LinearisedEuler hype.h:38 */
    /* point source and receiver of 3d/weno4/point_source (hype.h AppCtx isx..irz,
       rhs_function.c:337-346, source.c:11-17).  When source is not NULL, every
       RHS evaluation at time t adds source(t) / h^3 to component 1 of cell
       (isx, isy, isz); when receiver is not NULL, every RHS evaluation appends
       "t<TAB>u" (format "%e\t%.12e\n", component 0 at (irx, iry, irz) of the
       evaluated state) to it.  The fused TSSolve path does both per stage, with
       the stage times t, t + dt/2, t + dt, t + dt/2 of TSSSPRKS3. */
    PetscReal (*source)(PetscReal t);
    PetscInt isx, isy, isz;
    FILE *receiver;
    PetscInt irx, iry, irz;
} AppCtx;

/* Fills the operator fields with the reference's defaults (shipped linear
 * weights, c0 = rho0 = 1, periodic on all faces). */
PetscErrorCode AppCtxSetDefaults(AppCtx *ctx);

/* RHSFunction(TS, t, U, RHS, ctx): RHS = F(U) for the operator chosen in
 * ctx->system on the DMDA of the TS.  Same signature and registration as the
 * reference (hype.h:290, hype.c:171).  Errors: nonzero PetscErrorCode (the
 * reference's pressure-floor MPI_Abort, LinearisedEuler/pde.c:50-55, is not
 * reproduced: the linear operator has no such constraint). */
PetscErrorCode RHSFunction(TS ts, PetscReal t, Vec U, Vec RHS, void *ctx);

#ifdef __cplusplus
}
#endif
#endif
