/*
 * acoustic_ts.c -- TS paths of petsclite with the acoustic RHSFunction.
 *
 * Integrates the 3D linearised-acoustics system (LLF, CWENO4 with the given
 * linear weights) from seeded Gaussian pulses on an N^3 periodic unit cube
 * for `steps` steps and writes the final global vector (raw doubles) to OUT.
 *   mode fused   : TSSolve with RHSFunction registered (fused device stages)
 *   mode generic : TSSolve with a user RHS that calls RHSFunction (the
 *                  VecAXPY / VecAXPBYPCZ sequence of TSSSP RKS3)
 *   mode rk104   : generic path with TSSSPRK104
 * An optional last argument `nan` or `inf` puts that value into the pressure
 * of one cell of the initial state: TSSolve must then fail with the
 * non-finite-state error instead of returning a blown-up solution.
 *
 *   usage: acoustic_ts N steps gamma_hi gamma_lo mode OUT [nan|inf]
 */
#define nVar 4
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acoustic.h"
#include "acx.h"

static PetscErrorCode user_rhs(TS ts, PetscReal t, Vec U, Vec F, void *ctx)
{
    return RHSFunction(ts, t, U, F, ctx);
}

static PetscErrorCode run(int N, int steps, double ghi, double glo, const char *mode, const char *out,
                          const char *seed_bad)
{
    AppCtx c;
    DM da;
    Vec U;
    TS ts;
    AppCtxSetDefaults(&c);
    c.system = ACOUSTIC_SYSTEM;
    c.gammaHi = ghi;
    c.gammaLo = glo;
    c.N_x = c.N_y = c.N_z = N;
    c.x_min = c.y_min = c.z_min = 0.0;
    c.x_max = c.y_max = c.z_max = 1.0;
    c.h = 1.0 / N;
    c.CFL = 0.4;
    CHKERRQ(DMDACreate3d(PETSC_COMM_WORLD, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC, DM_BOUNDARY_PERIODIC,
                         DMDA_STENCIL_BOX, N, N, N, PETSC_DECIDE, PETSC_DECIDE, PETSC_DECIDE, nVar, 3, NULL, NULL,
                         NULL, &da));
    CHKERRQ(DMSetUp(da));
    CHKERRQ(DMCreateGlobalVector(da, &U));
    {
        acx_grid g;
        memset(&g, 0, sizeof(g));
        g.n[0] = g.n[1] = g.n[2] = N;
        g.system = ACX_ACOUSTIC4;
        g.h = c.h;
        g.gamma_hi = ghi;
        g.gamma_lo = glo;
        g.c0 = 1.0;
        g.rho0 = 1.0;
        double *host = (double *)malloc(sizeof(double) * 4 * (size_t)N * N * N);
        if (acx_field_gaussians(&g, 2025, 4, 1e-4, 1e-3, 0.02, 0.08, host)) return 90;
        Field ***u;
        CHKERRQ(DMDAVecGetArray(da, U, &u));
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i)
                    for (int v = 0; v < 4; ++v) u[k][j][i].comp[v] = host[(((size_t)k * N + j) * N + i) * 4 + v];
        if (seed_bad) u[N / 3][N / 2][N / 4].comp[0] = strcmp(seed_bad, "inf") == 0 ? INFINITY : NAN;
        CHKERRQ(DMDAVecRestoreArray(da, U, &u));
        free(host);
    }
    /* the first RHS call sets ctx->dt (LinearisedEuler/hype.c:170 + rhs_function.c:382-384) */
    Vec R;
    CHKERRQ(VecDuplicate(U, &R));
    CHKERRQ(TSCreate(PETSC_COMM_SELF, &ts));
    CHKERRQ(TSSetDM(ts, da));
    CHKERRQ(RHSFunction(ts, 0.0, U, R, &c));
    const int generic = strcmp(mode, "fused") != 0;
    CHKERRQ(TSSetRHSFunction(ts, NULL, generic ? user_rhs : RHSFunction, &c));
    CHKERRQ(TSSetTimeStep(ts, c.dt));
    CHKERRQ(TSSetMaxTime(ts, steps * c.dt));
    CHKERRQ(TSSetExactFinalTime(ts, TS_EXACTFINALTIME_MATCHSTEP));
    CHKERRQ(TSSetType(ts, TSSSP));
    CHKERRQ(TSSSPSetType(ts, strcmp(mode, "rk104") == 0 ? TSSSPRK104 : TSSSPRKS3));
    CHKERRQ(TSSolve(ts, U));
    PetscInt nsteps;
    CHKERRQ(TSGetStepNumber(ts, &nsteps));
    Field ***u;
    CHKERRQ(DMDAVecGetArrayRead(da, U, &u));
    FILE *f = fopen(out, "wb");
    if (!f) return 91;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j) fwrite(&u[k][j][0], sizeof(Field), N, f);
    fclose(f);
    CHKERRQ(DMDAVecRestoreArrayRead(da, U, &u));
    printf("steps %d dt %.17g\n", nsteps, c.dt);
    CHKERRQ(TSDestroy(&ts));
    CHKERRQ(VecDestroy(&U));
    CHKERRQ(VecDestroy(&R));
    CHKERRQ(DMDestroy(&da));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s N steps gamma_hi gamma_lo fused|generic|rk104 OUT\n", argv[0]);
        return 2;
    }
    PetscErrorCode e = PetscInitialize(&argc, &argv, NULL, NULL);
    if (!e) e = run(atoi(argv[1]), atoi(argv[2]), atof(argv[3]), atof(argv[4]), argv[5], argv[6],
                    argc > 7 ? argv[7] : NULL);
    if (e) {
        fprintf(stderr, "error %d: %s\n", e, PetscLiteLastError());
        return 1;
    }
    PetscFinalize();
    return 0;
}
