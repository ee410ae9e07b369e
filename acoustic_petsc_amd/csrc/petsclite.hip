// petsclite.hip -- one-rank DMDA / Vec / TS with device-resident vectors
// (include/petsclite.h) and the drop-in RHSFunction (include/acoustic.h).
//
// Semantics follow the PETSc calls made by the reference drivers
// (3d/weno4/plane_wave/hype.c:87-183, rhs_function.c:31-42):
//   * DMDA index space: owned box [0,M)x[0,N)x[0,P); ghost box widened by the
//     stencil width s on every axis that is periodic or ghosted.
//   * global -> local scatter: periodic axes wrap; on a ghosted axis the
//     ghost cells outside the domain keep what the local vector held
//     (user boundary code fills them, rhs_function.c:48-135).
//   * TSSSP RKS3 (s = n^2) and RK104 with the VecAXPY / VecAXPBYPCZ sequence
//     of PETSc's ssp.c; TS_EXACTFINALTIME_MATCHSTEP cuts the last step.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <sys/time.h>
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include "../../include/acoustic.h"
#include "../../include/acx.h"

static thread_local char g_perr[512] = "";

extern "C" PetscErrorCode PetscLiteError(PetscErrorCode code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_perr, sizeof(g_perr), fmt, ap);
    va_end(ap);
    return code;
}
extern "C" const char *PetscLiteLastError(void) { return g_perr; }

#define HIPCALL(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return PetscLiteError(76, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

// ---------------------------------------------------------------------------

struct _p_DM {
    int dim;
    int M[3];
    int dof, s;
    DMBoundaryType b[3];
    DMDAStencilType st;
    DM cda;          // coordinate DM (dof = dim)
    Vec coords;      // coordinate vector (global)
    bool is_coord;
    std::vector<Vec> local_pool;
    std::vector<std::string> fields;   // DMDASetFieldName
};

struct _p_Vec {
    DM dm;
    bool local;
    size_t n;             // scalars
    double *dev;
    double *host;
    int state;            // 0 both valid, 1 host newer, 2 device newer
    void *tables;         // pointer tables handed out by DMDAVecGetArray
    size_t tables_bytes;
};

struct _p_TS {
    DM dm;
    TSRHSFunction rhs;
    void *rhs_ctx;
    TSMonitorFunction mon;
    void *mon_ctx;
    double t, dt, max_time;
    int step;
    TSExactFinalTimeOption eft;
    int ssp;              // 1 rks3, 2 rk104
    int nstages;
    Vec w0, w1, F;
};

enum { VEC_BOTH = 0, VEC_HOST = 1, VEC_DEV = 2 };

static void box(DM da, bool local, int lo[3], int sz[3])
{
    for (int d = 0; d < 3; ++d) {
        lo[d] = 0;
        sz[d] = d < da->dim ? da->M[d] : 1;
        if (local && d < da->dim && da->b[d] != DM_BOUNDARY_NONE) {
            lo[d] = -da->s;
            sz[d] = da->M[d] + 2 * da->s;
        }
    }
}

static PetscErrorCode vec_create(DM da, bool local, Vec *out)
{
    int lo[3], sz[3];
    box(da, local, lo, sz);
    Vec v = new _p_Vec();
    v->dm = da;
    v->local = local;
    v->n = (size_t)sz[0] * sz[1] * sz[2] * da->dof;
    v->host = (double *)calloc(v->n, sizeof(double));
    if (!v->host) return PetscLiteError(55, "out of host memory");
    HIPCALL(hipMalloc(&v->dev, v->n * sizeof(double)));
    HIPCALL(hipMemset(v->dev, 0, v->n * sizeof(double)));
    v->state = VEC_BOTH;
    *out = v;
    return 0;
}

static PetscErrorCode to_host(Vec v)
{
    if (v->state == VEC_DEV) {
        HIPCALL(hipMemcpy(v->host, v->dev, v->n * sizeof(double), hipMemcpyDeviceToHost));
        v->state = VEC_BOTH;
    }
    return 0;
}

static PetscErrorCode to_dev(Vec v)
{
    if (v->state == VEC_HOST) {
        HIPCALL(hipMemcpy(v->dev, v->host, v->n * sizeof(double), hipMemcpyHostToDevice));
        v->state = VEC_BOTH;
    }
    return 0;
}

// ---- kernels --------------------------------------------------------------

__global__ void k_axpbypcz(size_t n, double *z, double a, double b, double c, const double *x, const double *y, int mode)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    // mode 0: z = a x + b y + c z; 1: z += a x (VecAXPY); 2: z = a x + b z (VecAXPBY); 3: z = a
    if (mode == 0) z[i] = a * x[i] + b * y[i] + c * z[i];
    else if (mode == 1) z[i] += a * x[i];
    else if (mode == 2) z[i] = a * x[i] + b * z[i];
    else z[i] = a;
}

// global -> local: every local cell whose index is owned after wrapping the
// periodic axes receives the owned value; other ghosts are left untouched.
__global__ void k_g2l(const double *g, double *l, int dof, int M0, int M1, int M2, int lo0, int lo1, int lo2,
                      int s0, int s1, int s2, int p0, int p1, int p2)
{
    size_t n = (size_t)s0 * s1 * s2;
    size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % s0) + lo0, j = (int)((t / s0) % s1) + lo1, k = (int)(t / ((size_t)s0 * s1)) + lo2;
    int ii = i, jj = j, kk = k;
    if (ii < 0 || ii >= M0) { if (!p0) return; ii = (ii + M0) % M0; }
    if (jj < 0 || jj >= M1) { if (!p1) return; jj = (jj + M1) % M1; }
    if (kk < 0 || kk >= M2) { if (!p2) return; kk = (kk + M2) % M2; }
    const double *src = g + (((size_t)kk * M1 + jj) * M0 + ii) * dof;
    double *dst = l + t * dof;
    for (int c = 0; c < dof; ++c) dst[c] = src[c];
}

static PetscErrorCode launch_axpbypcz(Vec z, double a, double b, double c, Vec x, Vec y, int mode)
{
    PetscErrorCode e;
    if ((e = to_dev(z))) return e;
    if (x && (e = to_dev(x))) return e;
    if (y && (e = to_dev(y))) return e;
    size_t n = z->n;
    if (n) {
        hipLaunchKernelGGL(k_axpbypcz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, z->dev, a, b, c,
                           x ? x->dev : nullptr, y ? y->dev : nullptr, mode);
        HIPCALL(hipGetLastError());
    }
    z->state = VEC_DEV;
    return 0;
}

// ---- runtime -------------------------------------------------------------

extern "C" {

PetscErrorCode PetscInitialize(int *argc, char ***argv, const char *file, const char *help)
{
    (void)argc; (void)argv; (void)file; (void)help;
    int n = 0;
    HIPCALL(hipGetDeviceCount(&n));
    if (n < 1) return PetscLiteError(76, "no GPU visible");
    return 0;
}

PetscErrorCode PetscFinalize(void)
{
    hipDeviceSynchronize();
    return 0;
}

PetscErrorCode PetscTime(PetscLogDouble *t)
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    *t = tv.tv_sec + 1e-6 * tv.tv_usec;
    return 0;
}

PetscErrorCode PetscPrintf(MPI_Comm comm, const char *fmt, ...)
{
    (void)comm;
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    return 0;
}

PetscErrorCode PetscFPrintf(MPI_Comm comm, FILE *fd, const char *fmt, ...)
{
    (void)comm;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(fd ? fd : stdout, fmt, ap);
    va_end(ap);
    return 0;
}

PetscErrorCode PetscObjectSetName(PetscObject obj, const char *name) { (void)obj; (void)name; return 0; }
MPI_Comm PetscObjectComm(PetscObject obj) { (void)obj; return PETSC_COMM_WORLD; }
PetscErrorCode MPI_Comm_size(MPI_Comm comm, PetscMPIInt *size) { (void)comm; *size = 1; return 0; }
PetscErrorCode MPI_Comm_rank(MPI_Comm comm, PetscMPIInt *rank) { (void)comm; *rank = 0; return 0; }

PetscErrorCode MPI_Allreduce(const void *in, void *out, int count, int type, int op, MPI_Comm comm)
{
    // one rank: the reduction of one contribution is the contribution
    (void)type; (void)op; (void)comm;
    memcpy(out, in, (size_t)count * sizeof(double));
    return 0;
}

// ---- DMDA ------------------------------------------------------------------

static PetscErrorCode dm_create(int dim, const DMBoundaryType *b, DMDAStencilType st, const PetscInt *M,
                                PetscInt dof, PetscInt s, DM *da)
{
    if (dof < 1 || s < 0) return PetscLiteError(62, "DMDA: dof >= 1 and stencil width >= 0 required");
    for (int d = 0; d < dim; ++d) {
        if (M[d] < 1) return PetscLiteError(62, "DMDA: axis %d has %d cells", d, M[d]);
        if (b[d] == DM_BOUNDARY_PERIODIC && M[d] < s)
            return PetscLiteError(62, "DMDA: periodic axis %d shorter than the stencil width", d);
    }
    DM dm = new _p_DM();
    dm->dim = dim;
    for (int d = 0; d < 3; ++d) {
        dm->M[d] = d < dim ? M[d] : 1;
        dm->b[d] = d < dim ? b[d] : DM_BOUNDARY_NONE;
    }
    dm->dof = dof;
    dm->s = s;
    dm->st = st;
    *da = dm;
    return 0;
}

PetscErrorCode DMDACreate3d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by, DMBoundaryType bz,
                            DMDAStencilType st, PetscInt M, PetscInt N, PetscInt P, PetscInt m, PetscInt n,
                            PetscInt p, PetscInt dof, PetscInt s, const PetscInt lx[], const PetscInt ly[],
                            const PetscInt lz[], DM *da)
{
    (void)comm; (void)lx; (void)ly; (void)lz;
    if ((m != PETSC_DECIDE && m != 1) || (n != PETSC_DECIDE && n != 1) || (p != PETSC_DECIDE && p != 1))
        return PetscLiteError(56, "DMDACreate3d: one rank only");
    DMBoundaryType b[3] = {bx, by, bz};
    PetscInt MM[3] = {M, N, P};
    return dm_create(3, b, st, MM, dof, s, da);
}

PetscErrorCode DMDACreate2d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by, DMDAStencilType st, PetscInt M,
                            PetscInt N, PetscInt m, PetscInt n, PetscInt dof, PetscInt s, const PetscInt lx[],
                            const PetscInt ly[], DM *da)
{
    (void)comm; (void)lx; (void)ly;
    if ((m != PETSC_DECIDE && m != 1) || (n != PETSC_DECIDE && n != 1))
        return PetscLiteError(56, "DMDACreate2d: one rank only");
    DMBoundaryType b[3] = {bx, by, DM_BOUNDARY_NONE};
    PetscInt MM[3] = {M, N, 1};
    return dm_create(2, b, st, MM, dof, s, da);
}

PetscErrorCode DMSetUp(DM da) { return da ? 0 : PetscLiteError(85, "DMSetUp: null DM"); }

PetscErrorCode DMDestroy(DM *da)
{
    if (!da || !*da) return 0;
    DM d = *da;
    for (Vec v : d->local_pool) VecDestroy(&v);
    if (d->coords) VecDestroy(&d->coords);
    if (d->cda) DMDestroy(&d->cda);
    delete d;
    *da = nullptr;
    return 0;
}

PetscErrorCode DMDASetFieldName(DM da, PetscInt nf, const char *name)
{
    if (nf < 0 || nf >= da->dof) return PetscLiteError(63, "field index out of range");
    if ((int)da->fields.size() < da->dof) da->fields.resize(da->dof);
    da->fields[nf] = name ? name : "";
    return 0;
}

PetscErrorCode DMDAGetCorners(DM da, PetscInt *xs, PetscInt *ys, PetscInt *zs, PetscInt *xm, PetscInt *ym,
                              PetscInt *zm)
{
    int lo[3], sz[3];
    box(da, false, lo, sz);
    if (xs) *xs = lo[0];
    if (ys) *ys = lo[1];
    if (zs) *zs = lo[2];
    if (xm) *xm = sz[0];
    if (ym) *ym = sz[1];
    if (zm) *zm = sz[2];
    return 0;
}

PetscErrorCode DMDAGetGhostCorners(DM da, PetscInt *xs, PetscInt *ys, PetscInt *zs, PetscInt *xm, PetscInt *ym,
                                   PetscInt *zm)
{
    int lo[3], sz[3];
    box(da, true, lo, sz);
    if (xs) *xs = lo[0];
    if (ys) *ys = lo[1];
    if (zs) *zs = lo[2];
    if (xm) *xm = sz[0];
    if (ym) *ym = sz[1];
    if (zm) *zm = sz[2];
    return 0;
}

PetscErrorCode DMCreateGlobalVector(DM da, Vec *v) { return vec_create(da, false, v); }
PetscErrorCode DMCreateLocalVector(DM da, Vec *v) { return vec_create(da, true, v); }

PetscErrorCode DMGetLocalVector(DM da, Vec *v)
{
    if (!da->local_pool.empty()) {
        *v = da->local_pool.back();
        da->local_pool.pop_back();
        return 0;
    }
    return vec_create(da, true, v);
}

PetscErrorCode DMRestoreLocalVector(DM da, Vec *v)
{
    da->local_pool.push_back(*v);
    *v = nullptr;
    return 0;
}

PetscErrorCode DMGlobalToLocalBegin(DM da, Vec g, InsertMode mode, Vec l)
{
    if (mode != INSERT_VALUES) return PetscLiteError(56, "DMGlobalToLocal: INSERT_VALUES only");
    if (g->local || !l->local || g->dm != da) return PetscLiteError(62, "DMGlobalToLocal: global and local vectors of this DM expected");
    PetscErrorCode e;
    if ((e = to_dev(g)) || (e = to_dev(l))) return e;
    int lo[3], sz[3];
    box(da, true, lo, sz);
    size_t n = (size_t)sz[0] * sz[1] * sz[2];
    hipLaunchKernelGGL(k_g2l, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, g->dev, l->dev, da->dof,
                       da->M[0], da->M[1], da->M[2], lo[0], lo[1], lo[2], sz[0], sz[1], sz[2],
                       da->b[0] == DM_BOUNDARY_PERIODIC, da->b[1] == DM_BOUNDARY_PERIODIC,
                       da->b[2] == DM_BOUNDARY_PERIODIC);
    HIPCALL(hipGetLastError());
    l->state = VEC_DEV;
    return 0;
}

PetscErrorCode DMGlobalToLocalEnd(DM da, Vec g, InsertMode mode, Vec l)
{
    (void)da; (void)g; (void)mode; (void)l;
    return 0;
}

// PETSc-style pointer tables: arr[k][j][i] (3D) or arr[j][i] (2D) addresses
// the dof-record of cell (i,j,k) in global indices, ghosts included for
// local vectors.  Rows are offset so that index lo maps to the first record.
static PetscErrorCode get_array(DM da, Vec v, void *array, bool read)
{
    PetscErrorCode e;
    if ((e = to_host(v))) return e;
    int lo[3], sz[3];
    box(da, v->local, lo, sz);
    const size_t rec = (size_t)da->dof;
    if (da->dim == 3) {
        size_t bytes = sizeof(double **) * sz[2] + sizeof(double *) * (size_t)sz[2] * sz[1];
        if (v->tables_bytes < bytes) {
            free(v->tables);
            v->tables = malloc(bytes);
            v->tables_bytes = bytes;
        }
        double ***planes = (double ***)v->tables;
        double **rows = (double **)(planes + sz[2]);
        for (int k = 0; k < sz[2]; ++k) {
            for (int j = 0; j < sz[1]; ++j)
                rows[(size_t)k * sz[1] + j] = v->host + (((size_t)k * sz[1] + j) * sz[0]) * rec - (ptrdiff_t)lo[0] * rec;
            planes[k] = rows + (size_t)k * sz[1] - lo[1];
        }
        *(double ****)array = planes - lo[2];
    } else if (da->dim == 2) {
        size_t bytes = sizeof(double *) * sz[1];
        if (v->tables_bytes < bytes) {
            free(v->tables);
            v->tables = malloc(bytes);
            v->tables_bytes = bytes;
        }
        double **rows = (double **)v->tables;
        for (int j = 0; j < sz[1]; ++j) rows[j] = v->host + ((size_t)j * sz[0]) * rec - (ptrdiff_t)lo[0] * rec;
        *(double ***)array = rows - lo[1];
    } else {
        *(double **)array = v->host - (ptrdiff_t)lo[0] * rec;
    }
    (void)read;
    return 0;
}

PetscErrorCode DMDAVecGetArray(DM da, Vec v, void *array) { return get_array(da, v, array, false); }
PetscErrorCode DMDAVecGetArrayRead(DM da, Vec v, void *array) { return get_array(da, v, array, true); }

PetscErrorCode DMDAVecRestoreArray(DM da, Vec v, void *array)
{
    (void)da; (void)array;
    v->state = VEC_HOST;   // the caller may have written through the array
    return 0;
}

PetscErrorCode DMDAVecRestoreArrayRead(DM da, Vec v, void *array)
{
    (void)da; (void)v; (void)array;
    return 0;
}

PetscErrorCode DMGetCoordinateDM(DM da, DM *cda)
{
    if (!da->cda) {
        DMBoundaryType b[3] = {da->b[0], da->b[1], da->b[2]};
        PetscInt M[3] = {da->M[0], da->M[1], da->M[2]};
        PetscErrorCode e = dm_create(da->dim, b, da->st, M, da->dim, da->s, &da->cda);
        if (e) return e;
        da->cda->is_coord = true;
    }
    *cda = da->cda;
    return 0;
}

PetscErrorCode DMGetCoordinates(DM da, Vec *c)
{
    if (!da->coords) return PetscLiteError(73, "DMGetCoordinates: call DMDASetUniformCoordinates first");
    *c = da->coords;
    return 0;
}

PetscErrorCode DMDASetUniformCoordinates(DM da, PetscReal xmin, PetscReal xmax, PetscReal ymin, PetscReal ymax,
                                         PetscReal zmin, PetscReal zmax)
{
    DM cda;
    PetscErrorCode e = DMGetCoordinateDM(da, &cda);
    if (e) return e;
    if (!da->coords && (e = vec_create(cda, false, &da->coords))) return e;
    const double mn[3] = {xmin, ymin, zmin}, mx[3] = {xmax, ymax, zmax};
    double hh[3];
    for (int d = 0; d < da->dim; ++d)
        hh[d] = da->b[d] == DM_BOUNDARY_PERIODIC ? (mx[d] - mn[d]) / da->M[d] : (mx[d] - mn[d]) / (da->M[d] - 1);
    Vec c = da->coords;
    if ((e = to_host(c))) return e;
    size_t cnt = 0;
    for (int k = 0; k < da->M[2]; ++k)
        for (int j = 0; j < da->M[1]; ++j)
            for (int i = 0; i < da->M[0]; ++i) {
                c->host[cnt++] = mn[0] + hh[0] * i;
                if (da->dim > 1) c->host[cnt++] = mn[1] + hh[1] * j;
                if (da->dim > 2) c->host[cnt++] = mn[2] + hh[2] * k;
            }
    c->state = VEC_HOST;
    return 0;
}

// ---- Vec ---------------------------------------------------------------------

PetscErrorCode VecDuplicate(Vec v, Vec *w)
{
    PetscErrorCode e = vec_create(v->dm->is_coord ? v->dm : v->dm, v->local, w);
    return e;
}

PetscErrorCode VecDestroy(Vec *v)
{
    if (!v || !*v) return 0;
    hipFree((*v)->dev);
    free((*v)->host);
    free((*v)->tables);
    delete *v;
    *v = nullptr;
    return 0;
}

PetscErrorCode VecGetSize(Vec v, PetscInt *n) { *n = (PetscInt)v->n; return 0; }
PetscErrorCode VecSet(Vec v, PetscScalar a) { return launch_axpbypcz(v, a, 0, 0, nullptr, nullptr, 3); }

PetscErrorCode VecCopy(Vec x, Vec y)
{
    if (x->n != y->n) return PetscLiteError(60, "VecCopy: size mismatch");
    PetscErrorCode e;
    if ((e = to_dev(x))) return e;
    const int rc = acx_vec_copy_f64(x->dev, y->dev, x->n, nullptr);
    if (rc) return PetscLiteError(76 + rc, "VecCopy: %s", acx_last_error());
    y->state = VEC_DEV;
    return 0;
}

PetscErrorCode VecAXPY(Vec y, PetscScalar a, Vec x) { return launch_axpbypcz(y, a, 0, 0, x, nullptr, 1); }
PetscErrorCode VecAXPBY(Vec y, PetscScalar a, PetscScalar b, Vec x) { return launch_axpbypcz(y, a, b, 0, x, nullptr, 2); }
PetscErrorCode VecAXPBYPCZ(Vec z, PetscScalar a, PetscScalar b, PetscScalar c, Vec x, Vec y)
{
    return launch_axpbypcz(z, a, b, c, x, y, 0);
}

PetscErrorCode VecNorm(Vec v, NormType t, PetscReal *nrm)
{
    PetscErrorCode e;
    if ((e = to_dev(v))) return e;
    double out[2];
    int rc = acx_norms_f64(v->dev, v->n, 1, -1, out, nullptr);
    if (rc) return PetscLiteError(76, "VecNorm: %s", acx_last_error());
    if (t == NORM_2 || t == NORM_FROBENIUS) *nrm = sqrt(out[0]);
    else if (t == NORM_INFINITY) *nrm = out[1];
    else return PetscLiteError(56, "VecNorm: NORM_2 and NORM_INFINITY only");
    return 0;
}

PetscErrorCode VecLiteGetDeviceArray(Vec v, double **d)
{
    PetscErrorCode e = to_dev(v);
    if (e) return e;
    v->state = VEC_DEV;   // the caller may write
    *d = v->dev;
    return 0;
}

// ---- Viewers ---------------------------------------------------------------

} // extern "C"

struct _p_PetscViewer {
    FILE *fp;
    bool binary;
    PetscFileMode mode;
    PetscViewerFormat fmt;
    std::string name;
    // VTK XML structured-grid viewer (PetscViewerVTKOpen): the grid and the
    // viewed vectors are collected and written when the viewer is destroyed
    bool vts = false;
    DM vts_dm = nullptr;
    std::vector<std::pair<std::string, std::vector<double>>> vts_arrays;
};

static void put_be32(FILE *fp, int32_t v)
{
    const uint32_t u = (uint32_t)v;
    const unsigned char b[4] = {(unsigned char)(u >> 24), (unsigned char)(u >> 16), (unsigned char)(u >> 8),
                                (unsigned char)u};
    fwrite(b, 1, 4, fp);
}

static bool get_be32(FILE *fp, int32_t *v)
{
    unsigned char b[4];
    if (fread(b, 1, 4, fp) != 4) return false;
    *v = (int32_t)(((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3]);
    return true;
}

static uint64_t bswap64(uint64_t x)
{
    x = ((x & 0x00000000FFFFFFFFull) << 32) | (x >> 32);
    x = ((x & 0x0000FFFF0000FFFFull) << 16) | ((x >> 16) & 0x0000FFFF0000FFFFull);
    return ((x & 0x00FF00FF00FF00FFull) << 8) | ((x >> 8) & 0x00FF00FF00FF00FFull);
}

static PetscErrorCode viewer_open(const char *name, bool binary, PetscFileMode mode, PetscViewer *out)
{
    if (!name || !out) return PetscLiteError(85, "viewer: null argument");
    const char *m = binary ? (mode == FILE_MODE_READ ? "rb" : mode == FILE_MODE_APPEND ? "ab" : "wb")
                           : (mode == FILE_MODE_APPEND ? "a" : "w");
    FILE *fp = fopen(name, m);
    if (!fp) return PetscLiteError(65, "cannot open %s", name);
    PetscViewer v = new _p_PetscViewer();
    v->fp = fp;
    v->binary = binary;
    v->mode = mode;
    v->fmt = PETSC_VIEWER_DEFAULT;
    v->name = name;
    *out = v;
    return 0;
}

extern "C" {

PetscErrorCode PetscViewerBinaryOpen(MPI_Comm comm, const char *name, PetscFileMode mode, PetscViewer *viewer)
{
    (void)comm;
    return viewer_open(name, true, mode, viewer);
}

PetscErrorCode PetscViewerASCIIOpen(MPI_Comm comm, const char *name, PetscViewer *viewer)
{
    (void)comm;
    return viewer_open(name, false, FILE_MODE_WRITE, viewer);
}

PetscErrorCode PetscViewerPushFormat(PetscViewer viewer, PetscViewerFormat format)
{
    if (!viewer) return PetscLiteError(85, "viewer: null");
    viewer->fmt = format;
    return 0;
}

PetscErrorCode PetscViewerPopFormat(PetscViewer viewer)
{
    if (!viewer) return PetscLiteError(85, "viewer: null");
    viewer->fmt = PETSC_VIEWER_DEFAULT;
    return 0;
}

PetscErrorCode PetscViewerVTKOpen(MPI_Comm comm, const char *name, PetscFileMode mode, PetscViewer *viewer)
{
    (void)comm;
    if (mode != FILE_MODE_WRITE) return PetscLiteError(56, "PetscViewerVTKOpen: FILE_MODE_WRITE only");
    PetscErrorCode e = viewer_open(name, true, mode, viewer);
    if (e) return e;
    (*viewer)->vts = true;
    return 0;
}

// VTK XML StructuredGrid (.vts), appended raw little-endian data with UInt64
// block headers: Position (3 x Float64 per point; DMDA coordinates when set,
// else the index), then one Float64 point array per viewed vector component.
static PetscErrorCode write_vts(PetscViewer v)
{
    DM da = v->vts_dm;
    if (!da && v->vts_arrays.empty()) return 0;
    if (!da) return PetscLiteError(73, "%s: VecView on a VTK viewer needs a DMView first", v->name.c_str());
    const int nx = da->M[0], ny = da->dim > 1 ? da->M[1] : 1, nz = da->dim > 2 ? da->M[2] : 1;
    const size_t np = (size_t)nx * ny * nz;
    std::vector<double> pos(3 * np, 0.0);
    const double *c = nullptr;
    if (da->coords) {
        PetscErrorCode e = to_host(da->coords);
        if (e) return e;
        c = da->coords->host;
    }
    for (size_t q = 0; q < np; ++q) {
        pos[3 * q] = (double)(q % nx);
        pos[3 * q + 1] = (double)((q / nx) % ny);
        pos[3 * q + 2] = (double)(q / ((size_t)nx * ny));
        if (c)
            for (int d = 0; d < da->dim; ++d) pos[3 * q + d] = c[q * da->dim + d];
    }
    FILE *fp = v->fp;
    const uint16_t probe = 1;
    if (*(const unsigned char *)&probe != 1) return PetscLiteError(56, "VTK writer: little-endian host required");
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
    fprintf(fp, "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", nx - 1, ny - 1, nz - 1);
    fprintf(fp, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", nx - 1, ny - 1, nz - 1);
    uint64_t off = 0;
    fprintf(fp, "      <Points>\n        <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" "
                "format=\"appended\" offset=\"%llu\"/>\n      </Points>\n", (unsigned long long)off);
    off += 8 + 8 * pos.size();
    fprintf(fp, "      <PointData>\n");
    for (const auto &a : v->vts_arrays) {
        fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\" format=\"appended\" "
                    "offset=\"%llu\"/>\n", a.first.c_str(), (unsigned long long)off);
        off += 8 + 8 * a.second.size();
    }
    fprintf(fp, "      </PointData>\n    </Piece>\n  </StructuredGrid>\n  <AppendedData encoding=\"raw\">\n_");
    auto block = [&](const std::vector<double> &d) {
        const uint64_t nb = 8 * d.size();
        return fwrite(&nb, 8, 1, fp) == 1 && fwrite(d.data(), 8, d.size(), fp) == d.size();
    };
    bool ok = block(pos);
    for (const auto &a : v->vts_arrays) ok = ok && block(a.second);
    fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
    return ok ? 0 : PetscLiteError(66, "%s: write failed", v->name.c_str());
}

PetscErrorCode PetscViewerDestroy(PetscViewer *viewer)
{
    if (!viewer || !*viewer) return 0;
    PetscErrorCode e = (*viewer)->vts ? write_vts(*viewer) : 0;
    if ((*viewer)->fp) fclose((*viewer)->fp);
    delete *viewer;
    *viewer = nullptr;
    return e;
}

PetscErrorCode DMView(DM da, PetscViewer viewer)
{
    if (!da || !viewer) return PetscLiteError(85, "DMView: null argument");
    if (viewer->vts) {
        viewer->vts_dm = da;
        return 0;
    }
    if (viewer->binary || viewer->fmt != PETSC_VIEWER_ASCII_VTK) {
        if (!viewer->binary)
            fprintf(viewer->fp, "DM Object: 1 MPI process\n  type: da\nDMDA %d x %d x %d, dof %d, stencil width %d\n",
                    da->M[0], da->dim > 1 ? da->M[1] : 1, da->dim > 2 ? da->M[2] : 1, da->dof, da->s);
        return 0;
    }
    const int nx = da->M[0], ny = da->dim > 1 ? da->M[1] : 1, nz = da->dim > 2 ? da->M[2] : 1;
    const size_t np = (size_t)nx * ny * nz;
    FILE *fp = viewer->fp;
    fprintf(fp, "# vtk DataFile Version 2.0\nGenerated by acoustic_petsc_amd petsclite\nASCII\n");
    fprintf(fp, "DATASET STRUCTURED_GRID\nDIMENSIONS %d %d %d\nPOINTS %zu double\n", nx, ny, nz, np);
    const double *c = nullptr;
    if (da->coords) {
        PetscErrorCode e = to_host(da->coords);
        if (e) return e;
        c = da->coords->host;
    }
    for (size_t q = 0; q < np; ++q) {
        const size_t i = q % nx, j = (q / nx) % ny, k = q / ((size_t)nx * ny);
        double p[3] = {(double)i, (double)j, (double)k};
        if (c)
            for (int d = 0; d < da->dim; ++d) p[d] = c[q * da->dim + d];
        fprintf(fp, "%.17g %.17g %.17g\n", p[0], p[1], p[2]);
    }
    fprintf(fp, "POINT_DATA %zu\n", np);
    return 0;
}

PetscErrorCode VecView(Vec v, PetscViewer viewer)
{
    if (!v || !viewer) return PetscLiteError(85, "VecView: null argument");
    if (v->local) return PetscLiteError(56, "VecView: global vectors only");
    PetscErrorCode e = to_host(v);
    if (e) return e;
    FILE *fp = viewer->fp;
    if (viewer->vts) {
        DM da = viewer->vts_dm ? viewer->vts_dm : v->dm;
        if (!da) return PetscLiteError(73, "VecView: VTK output needs a vector from a DMDA");
        const int dof = da->dof;
        const size_t np = v->n / dof;
        for (int f = 0; f < dof; ++f) {
            const bool named = f < (int)da->fields.size() && !da->fields[f].empty();
            std::vector<double> d(np);
            for (size_t q = 0; q < np; ++q) d[q] = v->host[q * dof + f];
            viewer->vts_arrays.emplace_back(named ? da->fields[f] : std::string("comp") + std::to_string(f),
                                            std::move(d));
        }
        if (!viewer->vts_dm) viewer->vts_dm = da;
        return 0;
    }
    if (viewer->binary) {
        if (viewer->mode == FILE_MODE_READ) return PetscLiteError(56, "VecView: viewer opened for reading");
        if (v->n > 0x7fffffff) return PetscLiteError(56, "VecView: vector longer than 2^31 - 1");
        put_be32(fp, VEC_FILE_CLASSID);
        put_be32(fp, (int32_t)v->n);
        std::vector<uint64_t> buf(v->n);
        for (size_t q = 0; q < v->n; ++q) {
            uint64_t u;
            memcpy(&u, &v->host[q], 8);
            buf[q] = bswap64(u);
        }
        if (fwrite(buf.data(), 8, v->n, fp) != v->n) return PetscLiteError(66, "VecView: write failed");
        fflush(fp);
        const int dof = v->dm ? v->dm->dof : 1;
        if (dof > 1) {
            FILE *info = fopen((viewer->name + ".info").c_str(), "w");
            if (info) {
                fprintf(info, "-vecload_block_size %d\n", dof);
                fclose(info);
            }
        }
        return 0;
    }
    if (viewer->fmt == PETSC_VIEWER_ASCII_VTK && v->dm) {
        const int dof = v->dm->dof;
        std::string name;
        for (int f = 0; f < dof; ++f) {
            const bool named = f < (int)v->dm->fields.size() && !v->dm->fields[f].empty();
            name += (f ? "_" : "") + (named ? v->dm->fields[f] : std::string("comp") + std::to_string(f));
        }
        fprintf(fp, "SCALARS %s double %d\nLOOKUP_TABLE default\n", name.c_str(), dof);
        const size_t np = v->n / dof;
        for (size_t q = 0; q < np; ++q) {
            for (int f = 0; f < dof; ++f) fprintf(fp, f ? " %.17g" : "%.17g", v->host[q * dof + f]);
            fputc('\n', fp);
        }
        return 0;
    }
    fprintf(fp, "Vec Object: 1 MPI process\n  type: seq\n");
    for (size_t q = 0; q < v->n; ++q) fprintf(fp, "%.17g\n", v->host[q]);
    return 0;
}

PetscErrorCode VecLoad(Vec v, PetscViewer viewer)
{
    if (!v || !viewer) return PetscLiteError(85, "VecLoad: null argument");
    if (!viewer->binary || viewer->mode != FILE_MODE_READ)
        return PetscLiteError(56, "VecLoad: needs a binary viewer opened with FILE_MODE_READ");
    if (v->local) return PetscLiteError(56, "VecLoad: global vectors only");
    int32_t cls = 0, n = 0;
    if (!get_be32(viewer->fp, &cls) || !get_be32(viewer->fp, &n))
        return PetscLiteError(66, "VecLoad: %s: truncated header", viewer->name.c_str());
    if (cls != VEC_FILE_CLASSID)
        return PetscLiteError(79, "VecLoad: %s: not a Vec record (class id %d)", viewer->name.c_str(), cls);
    if ((size_t)n != v->n)
        return PetscLiteError(60, "VecLoad: %s holds %d values, vector has %zu", viewer->name.c_str(), n, v->n);
    std::vector<uint64_t> buf(v->n);
    if (fread(buf.data(), 8, v->n, viewer->fp) != v->n)
        return PetscLiteError(66, "VecLoad: %s: truncated data", viewer->name.c_str());
    for (size_t q = 0; q < v->n; ++q) {
        const uint64_t u = bswap64(buf[q]);
        memcpy(&v->host[q], &u, 8);
    }
    v->state = VEC_HOST;
    return to_dev(v);
}

// ---- TS ------------------------------------------------------------------------

PetscErrorCode TSCreate(MPI_Comm comm, TS *ts)
{
    (void)comm;
    TS t = new _p_TS();
    t->ssp = 1;
    t->nstages = 4;
    t->max_time = 1e300;
    t->eft = TS_EXACTFINALTIME_UNSPECIFIED;
    *ts = t;
    return 0;
}

PetscErrorCode TSDestroy(TS *ts)
{
    if (!ts || !*ts) return 0;
    VecDestroy(&(*ts)->w0);
    VecDestroy(&(*ts)->w1);
    VecDestroy(&(*ts)->F);
    delete *ts;
    *ts = nullptr;
    return 0;
}

PetscErrorCode TSSetProblemType(TS ts, TSProblemType t) { (void)ts; (void)t; return 0; }
PetscErrorCode TSSetDM(TS ts, DM da) { ts->dm = da; return 0; }
PetscErrorCode TSGetDM(TS ts, DM *da) { *da = ts->dm; return ts->dm ? 0 : PetscLiteError(73, "TSGetDM: no DM set"); }
PetscErrorCode TSSetRHSFunction(TS ts, Vec r, TSRHSFunction f, void *ctx) { (void)r; ts->rhs = f; ts->rhs_ctx = ctx; return 0; }
PetscErrorCode TSSetStepNumber(TS ts, PetscInt step) { ts->step = step; return 0; }
PetscErrorCode TSGetStepNumber(TS ts, PetscInt *step) { *step = ts->step; return 0; }
PetscErrorCode TSSetTime(TS ts, PetscReal t) { ts->t = t; return 0; }
PetscErrorCode TSGetTime(TS ts, PetscReal *t) { *t = ts->t; return 0; }
PetscErrorCode TSSetTimeStep(TS ts, PetscReal dt) { ts->dt = dt; return 0; }
PetscErrorCode TSGetTimeStep(TS ts, PetscReal *dt) { *dt = ts->dt; return 0; }
PetscErrorCode TSMonitorSet(TS ts, TSMonitorFunction f, void *ctx, void (*destroy)(void **))
{
    (void)destroy;
    ts->mon = f;
    ts->mon_ctx = ctx;
    return 0;
}
PetscErrorCode TSSetMaxTime(TS ts, PetscReal t) { ts->max_time = t; return 0; }
PetscErrorCode TSSetExactFinalTime(TS ts, TSExactFinalTimeOption o) { ts->eft = o; return 0; }
PetscErrorCode TSSetFromOptions(TS ts) { (void)ts; return 0; }

PetscErrorCode TSSetType(TS ts, TSType type)
{
    (void)ts;
    return strcmp(type, TSSSP) == 0 ? 0 : PetscLiteError(86, "TSSetType: only %s is provided", TSSSP);
}

PetscErrorCode TSSSPSetType(TS ts, TSSSPType type)
{
    if (!strcmp(type, TSSSPRKS3)) ts->ssp = 1;
    else if (!strcmp(type, TSSSPRK104)) ts->ssp = 2;
    else return PetscLiteError(86, "TSSSPSetType: unknown type %s", type);
    return 0;
}

PetscErrorCode TSSSPSetNumStages(TS ts, PetscInt n)
{
    if (n < 1) return PetscLiteError(62, "TSSSPSetNumStages: at least one stage");
    ts->nstages = n;
    return 0;
}

} // extern "C"

// Is the registered RHS the library's own operator, so that whole stages can
// run fused on the device?
static bool own_operator(TS ts);
static PetscErrorCode fused_rks3_step(TS ts, Vec u, double dt, bool *done);

static PetscErrorCode rhs_eval(TS ts, double t, Vec q, Vec F) { return ts->rhs(ts, t, q, F, ts->rhs_ctx); }

// PETSc ssp.c sequence for SSPRK(s,3), s = n^2 (generic RHS)
static PetscErrorCode ssp_rks3(TS ts, double t0, double dt, Vec sol)
{
    const int s = ts->nstages;
    const int n = (int)(sqrt((double)s) + 0.001), r = s - n;
    if (n * n != s) return PetscLiteError(62, "TSSSPRKS3: the number of stages must be a square");
    PetscErrorCode e;
    auto frac = [&](int i) { return (i < n * (n + 1) / 2) ? 1. * i / (s - n) : (1. * i - n) / (s - n); };
    if ((e = VecCopy(sol, ts->w0))) return e;
    int i = 0;
    for (; i < (n - 1) * (n - 2) / 2; ++i) {
        if ((e = rhs_eval(ts, t0 + frac(i) * dt, ts->w0, ts->F))) return e;
        if ((e = VecAXPY(ts->w0, dt / r, ts->F))) return e;
    }
    if ((e = VecCopy(ts->w0, ts->w1))) return e;
    for (; i < n * (n + 1) / 2 - 1; ++i) {
        if ((e = rhs_eval(ts, t0 + frac(i) * dt, ts->w0, ts->F))) return e;
        if ((e = VecAXPY(ts->w0, dt / r, ts->F))) return e;
    }
    if ((e = rhs_eval(ts, t0 + frac(i) * dt, ts->w0, ts->F))) return e;
    if ((e = VecAXPBYPCZ(ts->w0, 1. * n / (2 * n - 1), (n - 1.) * dt / (r * (2 * n - 1)), (n - 1.) * 1. / (2 * n - 1),
                         ts->w1, ts->F)))
        return e;
    ++i;
    for (; i < s; ++i) {
        if ((e = rhs_eval(ts, t0 + frac(i) * dt, ts->w0, ts->F))) return e;
        if ((e = VecAXPY(ts->w0, dt / r, ts->F))) return e;
    }
    return VecCopy(ts->w0, sol);
}

// SSPRK(10,4)
static PetscErrorCode ssp_rk104(TS ts, double t0, double dt, Vec sol)
{
    static const double c[10] = {0, 1. / 6, 2. / 6, 3. / 6, 4. / 6, 2. / 6, 3. / 6, 4. / 6, 5. / 6, 1};
    PetscErrorCode e;
    int i;
    if ((e = VecCopy(sol, ts->w0))) return e;
    for (i = 0; i < 5; ++i) {
        if ((e = rhs_eval(ts, t0 + c[i] * dt, ts->w0, ts->F))) return e;
        if ((e = VecAXPY(ts->w0, dt / 6, ts->F))) return e;
    }
    if ((e = launch_axpbypcz(ts->w1, 1. / 25, 9. / 25, 0, sol, ts->w0, 0))) return e;   // gamma = 0: w1 not read
    if ((e = VecAXPBY(ts->w0, 15, -5, ts->w1))) return e;
    for (; i < 9; ++i) {
        if ((e = rhs_eval(ts, t0 + c[i] * dt, ts->w0, ts->F))) return e;
        if ((e = VecAXPY(ts->w0, dt / 6, ts->F))) return e;
    }
    if ((e = rhs_eval(ts, t0 + dt, ts->w0, ts->F))) return e;
    if ((e = VecAXPBYPCZ(ts->w1, 3. / 5, dt / 10, 1, ts->w0, ts->F))) return e;
    return VecCopy(ts->w1, sol);
}

extern "C" PetscErrorCode TSSolve(TS ts, Vec u)
{
    if (!ts->rhs) return PetscLiteError(73, "TSSolve: no RHS function");
    if (!(ts->dt > 0)) return PetscLiteError(62, "TSSolve: time step must be positive");
    PetscErrorCode e;
    if (!ts->w0) {
        if ((e = VecDuplicate(u, &ts->w0)) || (e = VecDuplicate(u, &ts->w1)) || (e = VecDuplicate(u, &ts->F))) return e;
    }
    const bool match = ts->eft == TS_EXACTFINALTIME_MATCHSTEP;
    if (ts->mon && (e = ts->mon(ts, ts->step, ts->t, u, ts->mon_ctx))) return e;
    double h = ts->dt;
    if (match && ts->t + h > ts->max_time) h = ts->max_time - ts->t;
    while (ts->t < ts->max_time) {
        bool done = false;
        if (ts->ssp == 1 && ts->nstages == 4 && own_operator(ts)) {
            if ((e = fused_rks3_step(ts, u, h, &done))) return e;
        }
        if (!done) {
            e = ts->ssp == 1 ? ssp_rks3(ts, ts->t, h, u) : ssp_rk104(ts, ts->t, h, u);
            if (e) return e;
        }
        ts->t += h;
        ts->step += 1;
        if (ts->mon && (e = ts->mon(ts, ts->step, ts->t, u, ts->mon_ctx))) return e;
        if (match && ts->t + h > ts->max_time) {
            const double next = ts->max_time - ts->t;
            if (next > PETSC_SMALL) h = next;
            else break;
        }
    }
    // a NaN or Inf written by any stage of the solve (the reference aborts on
    // an inadmissible state, 3d/weno3/LinearisedEuler/pde.c:50-55)
    const int rc = acx_nonfinite_check(nullptr);
    if (rc) return PetscLiteError(76 + rc, "TSSolve: %s", acx_last_error());
    return 0;
}

// ---- the drop-in right-hand side --------------------------------------------

static PetscErrorCode grid_from_ctx(DM da, const AppCtx *c, acx_grid *g)
{
    if (da->dim != 3 && da->dim != 2) return PetscLiteError(56, "RHSFunction: 2D or 3D DMDA expected");
    if (da->s < 3) return PetscLiteError(62, "RHSFunction: the DMDA stencil width must be at least 3 (hype.c:104)");
    memset(g, 0, sizeof(*g));
    // a 2D DMDA (2d/weno4/standing_wave/hype.c:94-107) is the one-plane grid n[2] = 1, periodic in z
    for (int d = 0; d < 3; ++d) g->n[d] = d < da->dim ? da->M[d] : 1;
    if (c->N_x != da->M[0] || c->N_y != da->M[1] || (da->dim == 3 && c->N_z != da->M[2]))
        return PetscLiteError(60, "RHSFunction: AppCtx N_x/N_y/N_z differ from the DMDA");
    g->system = c->system == ACOUSTIC_SYSTEM ? ACX_ACOUSTIC4 : ACX_WAVE2;
    const int want = g->system == ACX_ACOUSTIC4 ? 4 : 2;
    if (da->dof != want) return PetscLiteError(60, "RHSFunction: DMDA dof %d, operator needs %d", da->dof, want);
    const enum bndry_type lohi[3][2] = {{c->left_boundary, c->right_boundary},
                                        {c->bottom_boundary, c->top_boundary},
                                        {c->back_boundary, c->front_boundary}};
    for (int d = 0; d < 3; ++d) {
        const bool per = d >= da->dim || da->b[d] == DM_BOUNDARY_PERIODIC;
        if (per) {
            g->bc[d] = ACX_PERIODIC;
            continue;
        }
        if (lohi[d][0] != lohi[d][1] || lohi[d][0] == periodic)
            return PetscLiteError(56, "RHSFunction: non-periodic axis %d needs the same wall kind on both sides", d);
        g->bc[d] = lohi[d][0] == adiabatic ? ACX_ADIABATIC : lohi[d][0] == reflective ? ACX_REFLECTIVE : ACX_TRANSMISSIVE;
    }
    g->h = c->h;
    g->gamma_hi = c->gammaHi;
    g->gamma_lo = c->gammaLo;
    g->c0 = c->wave_speed;
    g->rho0 = c->rho_0;
    return 0;
}

extern "C" PetscErrorCode AppCtxSetDefaults(AppCtx *c)
{
    c->system = ACOUSTIC_SYSTEM;
    c->gammaHi = 1.0;
    c->gammaLo = 0.0;
    c->wave_speed = 1.0;
    c->rho_0 = 1.0;
    c->left_boundary = c->right_boundary = c->top_boundary = periodic;
    c->bottom_boundary = c->front_boundary = c->back_boundary = periodic;
    c->source = NULL;
    c->isx = c->isy = c->isz = 0;
    c->receiver = NULL;
    c->irx = c->iry = c->irz = 0;
    return 0;
}

// receiver sample and point source of one RHS evaluation at time t
// (3d/weno4/point_source/rhs_function.c:337-346); `scale` multiplies the
// source (1 for an RHS, the stage's dt coefficient for a fused stage output).
static PetscErrorCode point_terms_receiver(const AppCtx *c, const acx_grid *g, const double *U, double t)
{
    if (!c->receiver) return 0;
    double val = 0.0;
    int rc = acx_point_get_f64(g, U, c->irx, c->iry, c->irz, 0, &val, nullptr);
    if (rc) return PetscLiteError(76 + rc, "receiver: %s", acx_last_error());
    fprintf(c->receiver, "%e\t%.12e\n", t, val);
    return 0;
}

static PetscErrorCode point_terms_source(const AppCtx *c, const acx_grid *g, double *OUT, double t, double scale)
{
    if (!c->source) return 0;
    // Source(t) / h^3, or / h^2 on a 2D grid (2d/weno4/line_source/rhs_function.c:244-246)
    const double v = scale * c->source(t) / (g->h * g->h * (g->n[2] == 1 ? 1.0 : g->h));
    const int comp = g->system == ACX_WAVE2 ? 1 : 0;   // WAVE2: v_t; ACOUSTIC4: p_t
    int rc = acx_point_add_f64(g, OUT, c->isx, c->isy, c->isz, comp, v, nullptr);
    return rc ? PetscLiteError(76 + rc, "point source: %s", acx_last_error()) : 0;
}

extern "C" PetscErrorCode RHSFunction(TS ts, PetscReal t, Vec U, Vec RHS, void *ctx)
{
    AppCtx *c = (AppCtx *)ctx;
    DM da;
    PetscErrorCode e = TSGetDM(ts, &da);
    if (e) return e;
    acx_grid g;
    if ((e = grid_from_ctx(da, c, &g))) return e;
    double *du, *dr;
    if ((e = VecLiteGetDeviceArray(U, &du)) || (e = VecLiteGetDeviceArray(RHS, &dr))) return e;
    int rc = acx_rhs_f64(&g, du, dr, nullptr);
    if (rc) return PetscLiteError(76 + rc, "RHSFunction: %s", acx_last_error());
    if ((e = point_terms_source(c, &g, dr, t, 1.0))) return e;
    if ((e = point_terms_receiver(c, &g, du, t))) return e;
    if (g.system == ACX_ACOUSTIC4) {
        // LinearisedEuler/rhs_function.c:382-384 (one rank: the MIN-reduction is the value)
        c->dt = acx_cfl_dt(&g, c->CFL);
    }
    return 0;
}

static bool own_operator(TS ts) { return ts->rhs == RHSFunction && ts->dm && ts->dm->dim >= 2; }

static PetscErrorCode fused_rks3_step(TS ts, Vec u, double dt, bool *done)
{
    acx_grid g;
    PetscErrorCode e = grid_from_ctx(ts->dm, (const AppCtx *)ts->rhs_ctx, &g);
    if (e) return e;
    double *du, *d0, *d1;
    if ((e = VecLiteGetDeviceArray(u, &du)) || (e = VecLiteGetDeviceArray(ts->w0, &d0)) ||
        (e = VecLiteGetDeviceArray(ts->w1, &d1)))
        return e;
    const AppCtx *c = (const AppCtx *)ts->rhs_ctx;
    int rc;
    if (!c->source && !c->receiver) {
        rc = acx_ssp_rks3_step_f64(&g, du, d0, d1, dt, nullptr);
    } else {
        // the four stages of acx_ssp_rks3_step_f64 with the point terms of each
        // RHS evaluation (stage times t, t + dt/2, t + dt, t + dt/2)
        const double t = ts->t;
        struct { const double *q, *u0; double *out, a, b, c, ts; } st[4] = {
            {du, nullptr, d0, 0.0, 1.0, dt / 2, t},
            {d0, nullptr, d1, 0.0, 1.0, dt / 2, t + dt / 2},
            {d1, du, d0, 2.0 / 3.0, 1.0 / 3.0, dt / 6, t + dt},
            {d0, nullptr, du, 0.0, 1.0, dt / 2, t + dt / 2}};
        rc = 0;
        for (int k = 0; k < 4 && !rc; ++k) {
            if ((e = point_terms_receiver(c, &g, st[k].q, st[k].ts))) return e;
            rc = acx_stage_f64(&g, st[k].q, st[k].u0, st[k].out, st[k].a, st[k].b, st[k].c, nullptr);
            if (!rc && (e = point_terms_source(c, &g, st[k].out, st[k].ts, st[k].c))) return e;
        }
    }
    if (rc) return PetscLiteError(76 + rc, "TSSolve fused step: %s", acx_last_error());
    if (g.system == ACX_ACOUSTIC4) ((AppCtx *)ts->rhs_ctx)->dt = acx_cfl_dt(&g, ((AppCtx *)ts->rhs_ctx)->CFL);
    *done = true;
    return 0;
}
