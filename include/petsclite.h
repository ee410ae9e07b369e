/*
 * petsclite.h -- the PETSc surface the acoustic_petsc drivers call, on one
 * rank, with device-resident vectors on an MI355X.
 *
 * The reference programs build on PETSc DMDA / Vec / TS (3d/weno4/plane_wave/
 * hype.c:14-234, 3d/weno3/LinearisedEuler/hype.c, 2d/weno4/standing_wave/hype.c);
 * PETSc is not part of this build.  This header declares the subset those
 * drivers use with the same names, argument order, index conventions and
 * error-code behaviour, so a driver written against PETSc compiles against
 * it unchanged in the calls it makes:
 *
 *   - DMDA global vectors: owned cells, dof interlaced, x fastest;
 *     DMDAVecGetArray returns arrays indexed [k][j][i] (3D) / [j][i] (2D) in
 *     global indices; local vectors carry the stencil-width ghost box.
 *   - TSSSP with TSSSPRKS3 (s = n^2 stages, default 4) and TSSSPRK104;
 *     TS_EXACTFINALTIME_MATCHSTEP shortens the last step to land on max time.
 *   - Vec data live in HBM; DMDAVecGetArray / Restore synchronise a host
 *     mirror, device-side operations (TSSolve, VecAXPY, VecNorm, ...) never
 *     copy through the host.
 *
 * Every function returns a PetscErrorCode (0 = success); CHKERRQ propagates.
 */
#ifndef PETSCLITE_H
#define PETSCLITE_H

#include <stdio.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int PetscErrorCode;
typedef int PetscInt;
typedef int PetscMPIInt;
typedef double PetscReal;
typedef double PetscScalar;
typedef double PetscLogDouble;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef int MPI_Comm;

#define PETSC_COMM_WORLD ((MPI_Comm)0)
#define PETSC_COMM_SELF  ((MPI_Comm)1)
#define PETSC_DECIDE     (-1)
#define PETSC_PI         3.14159265358979323846264
#define PETSC_SMALL      1.0e-10
#define MPIU_REAL        0
#define MPIU_MIN         0
#define MPIU_MAX         1
#define MPIU_SUM         2

#define PetscMax(a, b) (((a) < (b)) ? (b) : (a))
#define PetscMin(a, b) (((a) < (b)) ? (a) : (b))
#define PetscPowReal(a, b) pow((a), (b))
#define PetscExpReal(a) exp(a)
#define PetscSqrtReal(a) sqrt(a)
#define PetscAbsReal(a) fabs(a)
#define PetscFunctionReturn(a) return (a)
#define CHKERRQ(e) do { PetscErrorCode e__ = (e); if (e__) return e__; } while (0)
#define SETERRQ(comm, code, msg) return PetscLiteError((code) ? (code) : 1, "%s", (msg))

typedef struct _p_DM *DM;
typedef struct _p_Vec *Vec;
typedef struct _p_TS *TS;
typedef struct _p_PetscViewer *PetscViewer;
typedef struct _p_PetscObject *PetscObject;

typedef enum { DM_BOUNDARY_NONE, DM_BOUNDARY_GHOSTED, DM_BOUNDARY_MIRROR, DM_BOUNDARY_PERIODIC } DMBoundaryType;
typedef enum { DMDA_STENCIL_STAR, DMDA_STENCIL_BOX } DMDAStencilType;
typedef enum { INSERT_VALUES, ADD_VALUES } InsertMode;
typedef enum { NORM_1 = 0, NORM_2 = 1, NORM_FROBENIUS = 2, NORM_INFINITY = 3 } NormType;
typedef enum { TS_LINEAR, TS_NONLINEAR } TSProblemType;
typedef enum { TS_EXACTFINALTIME_UNSPECIFIED, TS_EXACTFINALTIME_STEPOVER, TS_EXACTFINALTIME_INTERPOLATE,
               TS_EXACTFINALTIME_MATCHSTEP } TSExactFinalTimeOption;
typedef const char *TSType;
typedef const char *TSSSPType;
#define TSSSP      "ssp"
#define TSSSPRKS3  "rks3"
#define TSSSPRK104 "rk104"

typedef struct { PetscScalar x, y, z; } DMDACoor3d;
typedef struct { PetscScalar x, y; } DMDACoor2d;

typedef PetscErrorCode (*TSRHSFunction)(TS, PetscReal, Vec, Vec, void *);
typedef PetscErrorCode (*TSMonitorFunction)(TS, PetscInt, PetscReal, Vec, void *);

/* ---- runtime ------------------------------------------------------------- */
PetscErrorCode PetscInitialize(int *argc, char ***argv, const char *file, const char *help);
PetscErrorCode PetscFinalize(void);
PetscErrorCode PetscTime(PetscLogDouble *t);
PetscErrorCode PetscPrintf(MPI_Comm comm, const char *fmt, ...);
PetscErrorCode PetscFPrintf(MPI_Comm comm, FILE *fd, const char *fmt, ...);
PetscErrorCode PetscObjectSetName(PetscObject obj, const char *name);
MPI_Comm PetscObjectComm(PetscObject obj);
PetscErrorCode MPI_Comm_size(MPI_Comm comm, PetscMPIInt *size);
PetscErrorCode MPI_Comm_rank(MPI_Comm comm, PetscMPIInt *rank);
PetscErrorCode MPI_Allreduce(const void *in, void *out, int count, int type, int op, MPI_Comm comm);
/* last error message of the calling thread and the function that sets it */
const char *PetscLiteLastError(void);
PetscErrorCode PetscLiteError(PetscErrorCode code, const char *fmt, ...);

/* ---- DMDA ---------------------------------------------------------------- */
PetscErrorCode DMDACreate3d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by, DMBoundaryType bz,
                            DMDAStencilType st, PetscInt M, PetscInt N, PetscInt P, PetscInt m, PetscInt n,
                            PetscInt p, PetscInt dof, PetscInt s, const PetscInt lx[], const PetscInt ly[],
                            const PetscInt lz[], DM *da);
PetscErrorCode DMDACreate2d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by, DMDAStencilType st,
                            PetscInt M, PetscInt N, PetscInt m, PetscInt n, PetscInt dof, PetscInt s,
                            const PetscInt lx[], const PetscInt ly[], DM *da);
PetscErrorCode DMSetUp(DM da);
PetscErrorCode DMDestroy(DM *da);
PetscErrorCode DMDASetUniformCoordinates(DM da, PetscReal xmin, PetscReal xmax, PetscReal ymin, PetscReal ymax,
                                         PetscReal zmin, PetscReal zmax);
PetscErrorCode DMDASetFieldName(DM da, PetscInt nf, const char *name);
PetscErrorCode DMDAGetCorners(DM da, PetscInt *xs, PetscInt *ys, PetscInt *zs, PetscInt *xm, PetscInt *ym,
                              PetscInt *zm);
PetscErrorCode DMDAGetGhostCorners(DM da, PetscInt *xs, PetscInt *ys, PetscInt *zs, PetscInt *xm, PetscInt *ym,
                                   PetscInt *zm);
PetscErrorCode DMCreateGlobalVector(DM da, Vec *v);
PetscErrorCode DMCreateLocalVector(DM da, Vec *v);
PetscErrorCode DMGetLocalVector(DM da, Vec *v);
PetscErrorCode DMRestoreLocalVector(DM da, Vec *v);
PetscErrorCode DMGlobalToLocalBegin(DM da, Vec g, InsertMode mode, Vec l);
PetscErrorCode DMGlobalToLocalEnd(DM da, Vec g, InsertMode mode, Vec l);
PetscErrorCode DMDAVecGetArray(DM da, Vec v, void *array);
PetscErrorCode DMDAVecGetArrayRead(DM da, Vec v, void *array);
PetscErrorCode DMDAVecRestoreArray(DM da, Vec v, void *array);
PetscErrorCode DMDAVecRestoreArrayRead(DM da, Vec v, void *array);
PetscErrorCode DMGetCoordinateDM(DM da, DM *cda);
PetscErrorCode DMGetCoordinates(DM da, Vec *c);

/* ---- Vec ----------------------------------------------------------------- */
PetscErrorCode VecDuplicate(Vec v, Vec *w);
PetscErrorCode VecDestroy(Vec *v);
PetscErrorCode VecSet(Vec v, PetscScalar a);
PetscErrorCode VecCopy(Vec x, Vec y);
PetscErrorCode VecAXPY(Vec y, PetscScalar a, Vec x);
PetscErrorCode VecAXPBY(Vec y, PetscScalar a, PetscScalar b, Vec x);
PetscErrorCode VecAXPBYPCZ(Vec z, PetscScalar a, PetscScalar b, PetscScalar c, Vec x, Vec y);
PetscErrorCode VecNorm(Vec v, NormType t, PetscReal *nrm);
PetscErrorCode VecGetSize(Vec v, PetscInt *n);
/* device address of a vector's data (synchronised), for C-ABI calls */
PetscErrorCode VecLiteGetDeviceArray(Vec v, double **d);

/* ---- Viewers (3d/weno4/plane_wave/hype.c:147-151,190-196, monitor_function.c:22-58) ----
 * Binary: the PETSc binary format of VecView/VecLoad, one record per call:
 * int32 VEC_FILE_CLASSID (1211214), int32 length, then the values as float64,
 * all big-endian, in natural (x fastest, components interlaced) order; a
 * "<file>.info" text file carries "-vecload_block_size <dof>" for dof > 1.
 * ASCII + PETSC_VIEWER_ASCII_VTK: legacy VTK "STRUCTURED_GRID": DMView writes
 * the header and the point coordinates (cell centres from
 * DMDASetUniformCoordinates, else indices), VecView the POINT_DATA block with
 * the DMDA field name(s).  Plain ASCII: one value per line. */
typedef enum { FILE_MODE_READ, FILE_MODE_WRITE, FILE_MODE_APPEND } PetscFileMode;
typedef enum { PETSC_VIEWER_DEFAULT, PETSC_VIEWER_ASCII_VTK } PetscViewerFormat;
#define VEC_FILE_CLASSID 1211214
PetscErrorCode PetscViewerBinaryOpen(MPI_Comm comm, const char *name, PetscFileMode mode, PetscViewer *viewer);
PetscErrorCode PetscViewerASCIIOpen(MPI_Comm comm, const char *name, PetscViewer *viewer);
/* VTK XML structured grid (.vts): DMView records the grid, VecView the
 * vector's components; the file is written by PetscViewerDestroy
 * (3d/ctcs/point_source/wave.c:219-231). */
PetscErrorCode PetscViewerVTKOpen(MPI_Comm comm, const char *name, PetscFileMode mode, PetscViewer *viewer);
PetscErrorCode PetscViewerPushFormat(PetscViewer viewer, PetscViewerFormat format);
PetscErrorCode PetscViewerPopFormat(PetscViewer viewer);
PetscErrorCode PetscViewerDestroy(PetscViewer *viewer);
PetscErrorCode VecView(Vec v, PetscViewer viewer);
PetscErrorCode VecLoad(Vec v, PetscViewer viewer);
PetscErrorCode DMView(DM da, PetscViewer viewer);

/* ---- TS ------------------------------------------------------------------ */
PetscErrorCode TSCreate(MPI_Comm comm, TS *ts);
PetscErrorCode TSDestroy(TS *ts);
PetscErrorCode TSSetProblemType(TS ts, TSProblemType t);
PetscErrorCode TSSetDM(TS ts, DM da);
PetscErrorCode TSGetDM(TS ts, DM *da);
PetscErrorCode TSSetRHSFunction(TS ts, Vec r, TSRHSFunction f, void *ctx);
PetscErrorCode TSSetStepNumber(TS ts, PetscInt step);
PetscErrorCode TSGetStepNumber(TS ts, PetscInt *step);
PetscErrorCode TSSetTime(TS ts, PetscReal t);
PetscErrorCode TSGetTime(TS ts, PetscReal *t);
PetscErrorCode TSSetTimeStep(TS ts, PetscReal dt);
PetscErrorCode TSGetTimeStep(TS ts, PetscReal *dt);
PetscErrorCode TSMonitorSet(TS ts, TSMonitorFunction f, void *ctx, void (*destroy)(void **));
PetscErrorCode TSSetMaxTime(TS ts, PetscReal t);
PetscErrorCode TSSetExactFinalTime(TS ts, TSExactFinalTimeOption opt);
PetscErrorCode TSSetFromOptions(TS ts);
PetscErrorCode TSSetType(TS ts, TSType type);
PetscErrorCode TSSSPSetType(TS ts, TSSSPType type);
PetscErrorCode TSSSPSetNumStages(TS ts, PetscInt nstages);
PetscErrorCode TSSolve(TS ts, Vec u);

#ifdef __cplusplus
}
#endif
#endif
