/*
 * point_source_3d.c -- the reference WENO point-source program
 * (3d/weno4/point_source/hype.c:14-183, rhs_function.c:337-346, source.c:11-17)
 * on the PETSc-compatible surface: WAVE2 system (u, v), c = 250, periodic
 * cube of N cells of size h, zero initial state, Source(t) / h^3 added to v_t
 * at the source cell, receiver u written at every RHS evaluation, TSSSP RKS3
 * with 4 stages, fixed dt, nt steps.
 *
 *   usage: point_source_3d N h dt nt isx isy isz irx iry irz gamma_hi gamma_lo
 *                          fused|generic RECEIVER_OUT STATE_OUT
 */
#define nVar 2
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acoustic.h"

static PetscReal Source(PetscReal t)
{
    const PetscReal f0 = 100.0;       /* source.c:13 */
    const PetscReal t0 = 4.0 / f0;    /* source.c:14 */
    return -2.0 * (t - t0) * f0 * f0 * PetscExpReal(-1.0 * f0 * f0 * (t - t0) * (t - t0));
}

static PetscErrorCode user_rhs(TS ts, PetscReal t, Vec U, Vec F, void *ctx)
{
    return RHSFunction(ts, t, U, F, ctx);
}

int main(int argc, char **argv)
{
    if (argc < 16) {
        fprintf(stderr, "usage: %s N h dt nt isx isy isz irx iry irz gamma_hi gamma_lo fused|generic RCV STATE\n",
                argv[0]);
        return 2;
    }
    const int N = atoi(argv[1]);
    const double h = atof(argv[2]), dt = atof(argv[3]);
    const int nt = atoi(argv[4]);
    AppCtx Ctx;
    DM da;
    Vec U, RHS;
    TS ts;
    PetscErrorCode ierr = PetscInitialize(&argc, &argv, NULL, NULL);
    if (ierr) goto fail;
    AppCtxSetDefaults(&Ctx);
    Ctx.system = WAVE_SYSTEM;
    Ctx.wave_speed = 250.0;                      /* point_source/hype.h:39 */
    Ctx.gammaHi = atof(argv[11]);
    Ctx.gammaLo = atof(argv[12]);
    Ctx.N_x = Ctx.N_y = Ctx.N_z = N;
    Ctx.x_min = Ctx.y_min = Ctx.z_min = 0.0;
    Ctx.x_max = Ctx.y_max = Ctx.z_max = N * h;
    Ctx.h = h;
    Ctx.dt = dt;
    Ctx.CFL = Ctx.wave_speed * dt / h;           /* hype.c:44 */
    Ctx.source = Source;
    Ctx.isx = atoi(argv[5]); Ctx.isy = atoi(argv[6]); Ctx.isz = atoi(argv[7]);
    Ctx.irx = atoi(argv[8]); Ctx.iry = atoi(argv[9]); Ctx.irz = atoi(argv[10]);
    Ctx.receiver = fopen(argv[14], "w");
    if (!Ctx.receiver) { fprintf(stderr, "cannot open %s\n", argv[14]); return 1; }
    if ((ierr = DMDACreate3d(PETSC_COMM_WORLD, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC,
                             DMDA_STENCIL_BOX, N, N, N, PETSC_DECIDE, PETSC_DECIDE, PETSC_DECIDE, nVar, 3, NULL,
                             NULL, NULL, &da)))
        goto fail;
    if ((ierr = DMSetUp(da))) goto fail;
    if ((ierr = DMDASetFieldName(da, 0, "u")) || (ierr = DMDASetFieldName(da, 1, "v"))) goto fail;
    if ((ierr = DMCreateGlobalVector(da, &U)) || (ierr = VecDuplicate(U, &RHS))) goto fail;
    if ((ierr = VecSet(U, 0.0))) goto fail;      /* hype.c:154 */
    if ((ierr = TSCreate(PETSC_COMM_SELF, &ts)) || (ierr = TSSetProblemType(ts, TS_NONLINEAR)) ||
        (ierr = TSSetDM(ts, da)))
        goto fail;
    if ((ierr = RHSFunction(ts, 0.0, U, RHS, &Ctx))) goto fail;   /* hype.c:170 */
    if ((ierr = TSSetRHSFunction(ts, NULL, strcmp(argv[13], "fused") == 0 ? RHSFunction : user_rhs, &Ctx)))
        goto fail;
    if ((ierr = TSSetStepNumber(ts, 0)) || (ierr = TSSetTime(ts, 0.0)) || (ierr = TSSetTimeStep(ts, dt)) ||
        (ierr = TSSetMaxTime(ts, nt * dt)) || (ierr = TSSetExactFinalTime(ts, TS_EXACTFINALTIME_MATCHSTEP)) ||
        (ierr = TSSetType(ts, TSSSP)) || (ierr = TSSSPSetType(ts, TSSSPRKS3)) || (ierr = TSSSPSetNumStages(ts, 4)))
        goto fail;
    if ((ierr = TSSolve(ts, U))) goto fail;
    fclose(Ctx.receiver);
    {
        PetscViewer viewer;
        if ((ierr = PetscViewerBinaryOpen(PETSC_COMM_WORLD, argv[15], FILE_MODE_WRITE, &viewer)) ||
            (ierr = VecView(U, viewer)) || (ierr = PetscViewerDestroy(&viewer)))
            goto fail;
    }
    PetscInt steps;
    TSGetStepNumber(ts, &steps);
    printf("steps %d\n", steps);
    TSDestroy(&ts);
    VecDestroy(&U);
    VecDestroy(&RHS);
    DMDestroy(&da);
    PetscFinalize();
    return 0;
fail:
    fprintf(stderr, "error %d: %s\n", ierr, PetscLiteLastError());
    return 1;
}
