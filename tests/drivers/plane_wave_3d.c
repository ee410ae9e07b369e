/*
 * plane_wave_3d.c -- convergence driver for the 3D WAVE2 plane wave, written
 * against petsclite.h / acoustic.h the way the reference driver uses PETSc
 * (3d/weno4/plane_wave/hype.c:33-208): DMDA with a width-3 BOX stencil,
 * cell-centre coordinates, the initial and exact data integrated with the
 * reference's five-point rule and its corner offsets
 * (initialize_solution.c:41-54, error_norms.c:44-57), TSSSP RKS3 with 4
 * stages and MATCHSTEP, and the L2 (x h^1.5) / Linf error of component 0.
 *
 *   usage: plane_wave_3d N T      prints "N l2 linf steps"
 */
#define nVar 2
#include <math.h>
#include <stdlib.h>
#include "acoustic.h"

static const double gp5_x[5] = {0.0000000000000000, -0.2692346550528416, 0.2692346550528416,
                                -0.4530899229693320, 0.4530899229693320};
static const double gp5_w[5] = {0.28444444444444444, 0.23931433524968326, 0.23931433524968326,
                                0.11846344252809456, 0.11846344252809456};

static void wave(double x, double y, double z, double t, double q[2])
{
    const double k1 = 1.0, k2 = 1.0, k3 = 1.0, om = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
    q[0] = sin(k1 * x + k2 * y + k3 * z - om * t);
    q[1] = -om * cos(k1 * x + k2 * y + k3 * z - om * t);
}

/* cell integrals at time t into U (comp 1 zeroed when zero_v) */
static PetscErrorCode integrate(DM da, Vec U, double h, double t, int zero_v)
{
    DM cda;
    Vec cv;
    DMDACoor3d ***X;
    Field ***u;
    PetscInt xs, ys, zs, xm, ym, zm;
    CHKERRQ(DMDAGetCorners(da, &xs, &ys, &zs, &xm, &ym, &zm));
    CHKERRQ(DMGetCoordinateDM(da, &cda));
    CHKERRQ(DMGetCoordinates(da, &cv));
    CHKERRQ(DMDAVecGetArray(cda, cv, &X));
    CHKERRQ(DMDAVecGetArray(da, U, &u));
    for (PetscInt k = zs; k < zs + zm; ++k)
        for (PetscInt j = ys; j < ys + ym; ++j)
            for (PetscInt i = xs; i < xs + xm; ++i) {
                const double x0 = X[k][j][i].x - 0.5 * h, y0 = X[k][j][i].y - 0.5 * h, z0 = X[k][j][i].z - 0.5 * h;
                double acc[2] = {0.0, 0.0}, q[2];
                for (int l = 0; l < 5; l++)
                    for (int m = 0; m < 5; m++)
                        for (int n = 0; n < 5; ++n) {
                            wave(x0 + (double)l * h * gp5_x[l], y0 + (double)m * h * gp5_x[m],
                                 z0 + (double)n * h * gp5_x[n], t, q);
                            for (int c = 0; c < 2; ++c) acc[c] += gp5_w[l] * gp5_w[m] * gp5_w[n] * q[c];
                        }
                u[k][j][i].comp[0] = acc[0];
                u[k][j][i].comp[1] = zero_v ? 0.0 : acc[1];
            }
    CHKERRQ(DMDAVecRestoreArray(da, U, &u));
    CHKERRQ(DMDAVecRestoreArray(cda, cv, &X));
    return 0;
}

static PetscErrorCode zero_comp1(DM da, Vec U)
{
    Field ***u;
    PetscInt xs, ys, zs, xm, ym, zm;
    CHKERRQ(DMDAGetCorners(da, &xs, &ys, &zs, &xm, &ym, &zm));
    CHKERRQ(DMDAVecGetArray(da, U, &u));
    for (PetscInt k = zs; k < zs + zm; ++k)
        for (PetscInt j = ys; j < ys + ym; ++j)
            for (PetscInt i = xs; i < xs + xm; ++i) u[k][j][i].comp[1] = 0.0;
    CHKERRQ(DMDAVecRestoreArray(da, U, &u));
    return 0;
}

static PetscErrorCode run(int N, double T)
{
    AppCtx c;
    DM da;
    Vec U, E;
    TS ts;
    PetscInt steps;
    AppCtxSetDefaults(&c);
    c.system = WAVE_SYSTEM;
    c.x_min = c.y_min = c.z_min = 0.0;
    c.x_max = c.y_max = c.z_max = 2.0 * PETSC_PI;
    c.N_x = c.N_y = c.N_z = N;
    c.CFL = 0.5;
    c.h = (c.x_max - c.x_min) / (PetscReal)c.N_x;
    c.dt = (1. / 3.) * c.h * c.CFL / c.wave_speed;
    CHKERRQ(DMDACreate3d(PETSC_COMM_WORLD, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC,
                         DMDA_STENCIL_BOX, N, N, N, PETSC_DECIDE, PETSC_DECIDE, PETSC_DECIDE, nVar, 3, NULL, NULL,
                         NULL, &da));
    CHKERRQ(DMSetUp(da));
    CHKERRQ(DMCreateGlobalVector(da, &U));
    CHKERRQ(VecDuplicate(U, &E));
    CHKERRQ(DMDASetUniformCoordinates(da, c.x_min + 0.5 * c.h, c.x_max + 0.5 * c.h, c.y_min + 0.5 * c.h,
                                      c.y_max + 0.5 * c.h, c.z_min + 0.5 * c.h, c.z_max + 0.5 * c.h));
    CHKERRQ(integrate(da, U, c.h, 0.0, 0));
    CHKERRQ(TSCreate(PETSC_COMM_SELF, &ts));
    CHKERRQ(TSSetProblemType(ts, TS_NONLINEAR));
    CHKERRQ(TSSetDM(ts, da));
    CHKERRQ(TSSetRHSFunction(ts, NULL, RHSFunction, &c));
    CHKERRQ(TSSetStepNumber(ts, 0));
    CHKERRQ(TSSetTime(ts, 0.0));
    CHKERRQ(TSSetTimeStep(ts, c.dt));
    CHKERRQ(TSSetMaxTime(ts, T));
    CHKERRQ(TSSetExactFinalTime(ts, TS_EXACTFINALTIME_MATCHSTEP));
    CHKERRQ(TSSetType(ts, TSSSP));
    CHKERRQ(TSSSPSetType(ts, TSSSPRKS3));
    CHKERRQ(TSSSPSetNumStages(ts, 4));
    CHKERRQ(TSSolve(ts, U));
    CHKERRQ(TSGetStepNumber(ts, &steps));

    double l2, li;
    CHKERRQ(integrate(da, E, c.h, T, 1));
    CHKERRQ(zero_comp1(da, U));
    CHKERRQ(VecAXPY(E, -1.0, U));
    CHKERRQ(VecNorm(E, NORM_INFINITY, &li));
    CHKERRQ(VecNorm(E, NORM_2, &l2));
    l2 = l2 * pow(c.h, 1.5);
    PetscPrintf(PETSC_COMM_WORLD, "%d %.10e %.10e %d\n", N, l2, li, steps);
    CHKERRQ(TSDestroy(&ts));
    CHKERRQ(VecDestroy(&U));
    CHKERRQ(VecDestroy(&E));
    CHKERRQ(DMDestroy(&da));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s N T\n", argv[0]);
        return 2;
    }
    PetscErrorCode e = PetscInitialize(&argc, &argv, NULL, NULL);
    if (!e) e = run(atoi(argv[1]), atof(argv[2]));
    if (e) {
        fprintf(stderr, "error %d: %s\n", e, PetscLiteLastError());
        return 1;
    }
    PetscFinalize();
    return 0;
}
