/*
 * standing_wave_2d.c -- the reference 2D standing-wave program
 * (2d/weno4/standing_wave/hype.c:25-245) written against petsclite.h /
 * acoustic.h: WAVE2 on [-1, 1]^2 with adiabatic walls (DM_BOUNDARY_GHOSTED,
 * 2D DMDA, BOX stencil of width 3), cell-centre coordinates, initial data and
 * exact solution u = sin(2 pi x) sin(2 pi y) cos(omega t) integrated with the
 * centred five-point rule (initialize_solution.c:22-53, post_process.c:69-150),
 * dt = 0.5 h CFL / c with CFL 0.5, TSSSP RK104, MATCHSTEP to T = 1, and the
 * errors of component 0 (L2 / N_x, Linf) as Error.dat prints them.
 *
 *   usage: standing_wave_2d N      prints "N l2 linf steps"
 */
#define nVar 2
#include <math.h>
#include <stdlib.h>
#include "acoustic.h"

static const double gp5_x[5] = {0.0000000000000000, -0.2692346550528416, 0.2692346550528416,
                                -0.4530899229693320, 0.4530899229693320};
static const double gp5_w[5] = {0.28444444444444444, 0.23931433524968326, 0.23931433524968326,
                                0.11846344252809456, 0.11846344252809456};

/* initial_condition.c:12-36 */
static void exact(double x, double y, double t, double q[2])
{
    const double kx = 2 * PETSC_PI, ky = 2 * PETSC_PI, om = sqrt(kx * kx + ky * ky);
    q[0] = sin(kx * x) * sin(ky * y) * cos(om * t);
    q[1] = -om * sin(kx * x) * sin(ky * y) * sin(om * t);
}

/* cell averages at time t into U (comp 1 zeroed when zero_v) */
static PetscErrorCode integrate(DM da, Vec U, double h, double t, int zero_v)
{
    DM cda;
    Vec cv;
    DMDACoor2d **X;
    Field **u;
    PetscInt xs, ys, xm, ym;
    CHKERRQ(DMDAGetCorners(da, &xs, &ys, NULL, &xm, &ym, NULL));
    CHKERRQ(DMGetCoordinateDM(da, &cda));
    CHKERRQ(DMGetCoordinates(da, &cv));
    CHKERRQ(DMDAVecGetArray(cda, cv, &X));
    CHKERRQ(DMDAVecGetArray(da, U, &u));
    for (PetscInt j = ys; j < ys + ym; ++j)
        for (PetscInt i = xs; i < xs + xm; ++i) {
            double acc[2] = {0.0, 0.0}, q[2];
            for (int l = 0; l < 5; ++l)
                for (int m = 0; m < 5; ++m) {
                    exact(X[j][i].x + h * gp5_x[l], X[j][i].y + h * gp5_x[m], t, q);
                    for (int c = 0; c < 2; ++c) acc[c] += gp5_w[l] * gp5_w[m] * q[c];
                }
            u[j][i].comp[0] = acc[0];
            u[j][i].comp[1] = zero_v ? 0.0 : acc[1];
        }
    CHKERRQ(DMDAVecRestoreArray(da, U, &u));
    CHKERRQ(DMDAVecRestoreArray(cda, cv, &X));
    return 0;
}

static PetscErrorCode zero_comp1(DM da, Vec U)
{
    Field **u;
    PetscInt xs, ys, xm, ym;
    CHKERRQ(DMDAGetCorners(da, &xs, &ys, NULL, &xm, &ym, NULL));
    CHKERRQ(DMDAVecGetArray(da, U, &u));
    for (PetscInt j = ys; j < ys + ym; ++j)
        for (PetscInt i = xs; i < xs + xm; ++i) u[j][i].comp[1] = 0.0;
    CHKERRQ(DMDAVecRestoreArray(da, U, &u));
    return 0;
}

static PetscErrorCode run(int N)
{
    AppCtx c;
    DM da;
    Vec U, E;
    TS ts;
    PetscInt steps;
    AppCtxSetDefaults(&c);
    c.system = WAVE_SYSTEM;
    c.x_min = c.y_min = -1.0;
    c.x_max = c.y_max = 1.0;
    c.N_x = c.N_y = N;
    c.N_z = 1;
    c.CFL = 0.5;
    c.left_boundary = c.right_boundary = c.bottom_boundary = c.top_boundary = adiabatic;
    c.h = (c.x_max - c.x_min) / (PetscReal)c.N_x;
    c.dt = 0.5 * c.h * c.CFL / c.wave_speed;                     /* hype.c:49-50 */
    CHKERRQ(DMDACreate2d(PETSC_COMM_WORLD, DM_BOUNDARY_GHOSTED, DM_BOUNDARY_GHOSTED, DMDA_STENCIL_BOX, N, N,
                         PETSC_DECIDE, PETSC_DECIDE, nVar, 3, NULL, NULL, &da));
    CHKERRQ(DMSetUp(da));
    CHKERRQ(DMCreateGlobalVector(da, &U));
    CHKERRQ(VecDuplicate(U, &E));
    CHKERRQ(DMDASetUniformCoordinates(da, c.x_min + 0.5 * c.h, c.x_max - 0.5 * c.h, c.y_min + 0.5 * c.h,
                                      c.y_max - 0.5 * c.h, 0.0, 0.0));
    CHKERRQ(DMDASetFieldName(da, 0, "u"));
    CHKERRQ(DMDASetFieldName(da, 1, "v"));
    CHKERRQ(integrate(da, U, c.h, 0.0, 0));
    CHKERRQ(TSCreate(PETSC_COMM_SELF, &ts));
    CHKERRQ(TSSetProblemType(ts, TS_NONLINEAR));
    CHKERRQ(TSSetDM(ts, da));
    CHKERRQ(TSSetRHSFunction(ts, NULL, RHSFunction, &c));
    CHKERRQ(TSSetStepNumber(ts, 0));
    CHKERRQ(TSSetTime(ts, 0.0));
    CHKERRQ(TSSetTimeStep(ts, c.dt));
    CHKERRQ(TSSetMaxTime(ts, 1.0));
    CHKERRQ(TSSetExactFinalTime(ts, TS_EXACTFINALTIME_MATCHSTEP));
    CHKERRQ(TSSetType(ts, TSSSP));
    CHKERRQ(TSSSPSetType(ts, TSSSPRK104));
    CHKERRQ(TSSolve(ts, U));
    CHKERRQ(TSGetStepNumber(ts, &steps));

    double l2, li;
    CHKERRQ(integrate(da, E, c.h, 1.0, 1));
    CHKERRQ(zero_comp1(da, U));
    CHKERRQ(VecAXPY(E, -1.0, U));
    CHKERRQ(VecNorm(E, NORM_INFINITY, &li));
    CHKERRQ(VecNorm(E, NORM_2, &l2));
    l2 = l2 / (PetscReal)c.N_x;                                  /* post_process.c:141 */
    PetscPrintf(PETSC_COMM_WORLD, "%d %.10e %.10e %d\n", N, l2, li, steps);
    CHKERRQ(TSDestroy(&ts));
    CHKERRQ(VecDestroy(&U));
    CHKERRQ(VecDestroy(&E));
    CHKERRQ(DMDestroy(&da));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s N\n", argv[0]);
        return 2;
    }
    PetscErrorCode e = PetscInitialize(&argc, &argv, NULL, NULL);
    if (!e) e = run(atoi(argv[1]));
    if (e) {
        fprintf(stderr, "error %d: %s\n", e, PetscLiteLastError());
        return 1;
    }
    PetscFinalize();
    return 0;
}
