/*
 * restart_3d.c -- the output and restart pattern of the reference drivers
 * (3d/weno4/plane_wave/hype.c:115-125,144-160,188-196 and
 * monitor_function.c:12-62) on the PETSc-compatible surface.
 *
 * ACOUSTIC4, N^3 periodic unit cube, linear CWENO4, seeded Gaussian pulses.
 *   run    : TSSolve for `steps` steps; the monitor prints "step t = time",
 *            writes DIR/plot/sol-%05d.vtk every write_interval steps and
 *            DIR/restart1.bin every restart_interval steps (DIR/restart2.bin
 *            every restart_interval + 7 steps), as MonitorFunction does; at
 *            the end DIR/final.bin (VecView, binary) and
 *            DIR/plot/sol-%08d.vtk (DMView + VecView, ASCII VTK) and
 *            DIR/plot/sol-%08d.vts (PetscViewerVTKOpen, XML structured grid).
 *   resume : VecLoad of DIR/restart1.bin, which holds step
 *            k = restart_interval * floor(steps / restart_interval), then
 *            TSSolve from step k to step `steps`; writes DIR/final_resumed.bin.
 *
 *   usage: restart_3d N steps restart_interval write_interval run|resume DIR
 */
#define nVar 4
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acoustic.h"
#include "acx.h"

typedef struct {
    AppCtx app;
    int write_interval, restart_interval;
    const char *dir;
} RunCtx;

static PetscErrorCode monitor(TS ts, PetscInt step, PetscReal time, Vec U, void *ctx)
{
    RunCtx *r = (RunCtx *)ctx;
    char name[512];
    PetscViewer viewer;
    CHKERRQ(PetscPrintf(PETSC_COMM_WORLD, "%d t = %f\n", step, time));
    if (r->write_interval != 0 && step % r->write_interval == 0) {
        DM da;
        CHKERRQ(TSGetDM(ts, &da));
        snprintf(name, sizeof(name), "%s/plot/sol-%05d.vtk", r->dir, step);
        CHKERRQ(PetscViewerASCIIOpen(PETSC_COMM_WORLD, name, &viewer));
        CHKERRQ(PetscViewerPushFormat(viewer, PETSC_VIEWER_ASCII_VTK));
        CHKERRQ(DMView(da, viewer));
        CHKERRQ(VecView(U, viewer));
        CHKERRQ(PetscViewerDestroy(&viewer));
    }
    if (r->restart_interval != 0) {
        if (step % r->restart_interval == 0) {
            snprintf(name, sizeof(name), "%s/restart1.bin", r->dir);
            CHKERRQ(PetscViewerBinaryOpen(PETSC_COMM_WORLD, name, FILE_MODE_WRITE, &viewer));
            CHKERRQ(VecView(U, viewer));
            CHKERRQ(PetscViewerDestroy(&viewer));
        }
        if (step % (r->restart_interval + 7) == 0) {
            snprintf(name, sizeof(name), "%s/restart2.bin", r->dir);
            CHKERRQ(PetscViewerBinaryOpen(PETSC_COMM_WORLD, name, FILE_MODE_WRITE, &viewer));
            CHKERRQ(VecView(U, viewer));
            CHKERRQ(PetscViewerDestroy(&viewer));
        }
    }
    return 0;
}

static PetscErrorCode initial(DM da, Vec U, int N, double h)
{
    acx_grid g;
    memset(&g, 0, sizeof(g));
    g.n[0] = g.n[1] = g.n[2] = N;
    g.system = ACX_ACOUSTIC4;
    g.h = h;
    g.gamma_hi = 1.0;
    g.c0 = 1.0;
    g.rho0 = 1.0;
    double *host = (double *)malloc(sizeof(double) * 4 * (size_t)N * N * N);
    if (!host || acx_field_gaussians(&g, 77, 3, 1e-4, 1e-3, 0.05, 0.1, host)) return 90;
    Field ***u;
    CHKERRQ(DMDAVecGetArray(da, U, &u));
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i)
                for (int v = 0; v < 4; ++v) u[k][j][i].comp[v] = host[(((size_t)k * N + j) * N + i) * 4 + v];
    CHKERRQ(DMDAVecRestoreArray(da, U, &u));
    free(host);
    return 0;
}

static PetscErrorCode run(int N, int steps, int ri, int wi, int resume, const char *dir)
{
    RunCtx r;
    DM da;
    Vec U, R;
    TS ts;
    char name[512];
    PetscViewer viewer;
    AppCtxSetDefaults(&r.app);
    r.app.system = ACOUSTIC_SYSTEM;
    r.app.N_x = r.app.N_y = r.app.N_z = N;
    r.app.x_min = r.app.y_min = r.app.z_min = 0.0;
    r.app.x_max = r.app.y_max = r.app.z_max = 1.0;
    r.app.h = 1.0 / N;
    r.app.CFL = 0.4;
    r.write_interval = resume ? 0 : wi;
    r.restart_interval = resume ? 0 : ri;
    r.dir = dir;
    CHKERRQ(DMDACreate3d(PETSC_COMM_WORLD, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC,
                         DMDA_STENCIL_BOX, N, N, N, PETSC_DECIDE, PETSC_DECIDE, PETSC_DECIDE, nVar, 3, NULL, NULL,
                         NULL, &da));
    CHKERRQ(DMSetUp(da));
    CHKERRQ(DMDASetUniformCoordinates(da, 0.5 * r.app.h, 1.0 + 0.5 * r.app.h, 0.5 * r.app.h, 1.0 + 0.5 * r.app.h,
                                      0.5 * r.app.h, 1.0 + 0.5 * r.app.h));
    CHKERRQ(DMDASetFieldName(da, 0, "P"));
    CHKERRQ(DMDASetFieldName(da, 1, "U"));
    CHKERRQ(DMDASetFieldName(da, 2, "V"));
    CHKERRQ(DMDASetFieldName(da, 3, "W"));
    CHKERRQ(DMCreateGlobalVector(da, &U));
    CHKERRQ(VecDuplicate(U, &R));
    int k0 = 0;
    if (resume) {
        k0 = ri * (steps / ri);
        snprintf(name, sizeof(name), "%s/restart1.bin", dir);
        CHKERRQ(PetscPrintf(PETSC_COMM_WORLD, "reading vector in binary from %s ...\n", name));
        CHKERRQ(PetscViewerBinaryOpen(PETSC_COMM_WORLD, name, FILE_MODE_READ, &viewer));
        CHKERRQ(VecLoad(U, viewer));
        CHKERRQ(PetscViewerDestroy(&viewer));
    } else {
        CHKERRQ(initial(da, U, N, r.app.h));
    }
    CHKERRQ(TSCreate(PETSC_COMM_SELF, &ts));
    CHKERRQ(TSSetProblemType(ts, TS_NONLINEAR));
    CHKERRQ(TSSetDM(ts, da));
    CHKERRQ(RHSFunction(ts, 0.0, U, R, &r.app));   /* sets r.app.dt */
    CHKERRQ(TSSetRHSFunction(ts, NULL, RHSFunction, &r.app));
    /* step and time of the restart point, as a resumed reference run would set them */
    double t0 = 0.0;
    for (int s = 0; s < k0; ++s) t0 += r.app.dt;
    CHKERRQ(TSSetStepNumber(ts, k0));
    CHKERRQ(TSSetTime(ts, t0));
    CHKERRQ(TSSetTimeStep(ts, r.app.dt));
    CHKERRQ(TSMonitorSet(ts, monitor, &r, NULL));
    double tf = 0.0;
    for (int s = 0; s < steps; ++s) tf += r.app.dt;
    CHKERRQ(TSSetMaxTime(ts, tf));
    CHKERRQ(TSSetExactFinalTime(ts, TS_EXACTFINALTIME_MATCHSTEP));
    CHKERRQ(TSSetFromOptions(ts));
    CHKERRQ(TSSetType(ts, TSSSP));
    CHKERRQ(TSSSPSetType(ts, TSSSPRKS3));
    CHKERRQ(TSSSPSetNumStages(ts, 4));
    CHKERRQ(TSSolve(ts, U));
    PetscInt nsteps;
    CHKERRQ(TSGetStepNumber(ts, &nsteps));
    snprintf(name, sizeof(name), "%s/%s", dir, resume ? "final_resumed.bin" : "final.bin");
    CHKERRQ(PetscViewerBinaryOpen(PETSC_COMM_WORLD, name, FILE_MODE_WRITE, &viewer));
    CHKERRQ(VecView(U, viewer));
    CHKERRQ(PetscViewerDestroy(&viewer));
    if (!resume) {
        snprintf(name, sizeof(name), "%s/plot/sol-%08d.vtk", dir, nsteps);
        CHKERRQ(PetscViewerASCIIOpen(PETSC_COMM_WORLD, name, &viewer));
        CHKERRQ(PetscViewerPushFormat(viewer, PETSC_VIEWER_ASCII_VTK));
        CHKERRQ(DMView(da, viewer));
        CHKERRQ(VecView(U, viewer));
        CHKERRQ(PetscViewerDestroy(&viewer));
        /* XML structured grid, as write_vts of 3d/ctcs/point_source/wave.c:219-231 */
        snprintf(name, sizeof(name), "%s/plot/sol-%08d.vts", dir, nsteps);
        CHKERRQ(PetscViewerVTKOpen(PETSC_COMM_WORLD, name, FILE_MODE_WRITE, &viewer));
        CHKERRQ(DMView(da, viewer));
        CHKERRQ(VecView(U, viewer));
        CHKERRQ(PetscViewerDestroy(&viewer));
    }
    printf("steps %d dt %.17g\n", nsteps, r.app.dt);
    CHKERRQ(TSDestroy(&ts));
    CHKERRQ(VecDestroy(&U));
    CHKERRQ(VecDestroy(&R));
    CHKERRQ(DMDestroy(&da));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s N steps restart_interval write_interval run|resume DIR\n", argv[0]);
        return 2;
    }
    PetscErrorCode e = PetscInitialize(&argc, &argv, NULL, NULL);
    if (!e)
        e = run(atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), strcmp(argv[5], "resume") == 0, argv[6]);
    if (e) {
        fprintf(stderr, "error %d: %s\n", e, PetscLiteLastError());
        return 1;
    }
    PetscFinalize();
    return 0;
}
