// acx_api.hip -- the C ABI (include/acx.h): argument checks, kernel choice
// and the SSP stage sequences.
#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int env_int(const char *name, int dflt)
{
    const char *s = getenv(name);
    return (s && *s) ? atoi(s) : dflt;
}

// Automatic z-chunk of a 2.5D march.  A launch runs in ceil(workgroups /
// slots) rounds (slots = CUs x workgroups that fit a CU), each about as long
// as one workgroup's march of zc planes plus a fixed start of `start` planes'
// worth (z-halo planes, prologue); the zc of 8 .. 128 that minimises
// rounds x (zc + start) is taken, the larger on a tie.
int auto_zc(int ntx, int nty, int nz, int slots, int start)
{
    static const int cand[] = {8, 16, 32, 48, 64, 96, 128};
    int best = nz < 8 ? nz : 8;
    double best_cost = -1.0;
    for (int zc : cand) {
        if (zc > nz && zc != 8) break;
        const int z = zc < nz ? zc : nz;
        const long long wg = (long long)ntx * nty * ((nz + z - 1) / z);
        const long long rounds = (wg + slots - 1) / slots;
        const double cost = (double)rounds * (z + start);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best = z;
        }
    }
    return best;
}

// CUs of the current device (cached per device).
int device_cus()
{
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return ACX_OK;
    return set_error(ACX_EHIP, "%s: %s", what, hipGetErrorString(e));
}


static int check_grid(const acx_grid *g)
{
    if (!g) return set_error(ACX_EINVAL, "grid is NULL");
    for (int d = 0; d < 3; ++d) {
        if (g->bc[d] < ACX_PERIODIC || g->bc[d] > ACX_TRANSMISSIVE)
            return set_error(ACX_EINVAL, "bc[%d] = %d is not a boundary kind", d, g->bc[d]);
        // n[2] = 1 with a periodic z axis is a 2D grid (one plane)
        if (d == 2 && g->n[2] == 1 && g->bc[2] == ACX_PERIODIC) continue;
        if (g->n[d] < 4) return set_error(ACX_EINVAL, "n[%d] = %d: at least 4 cells per axis are required", d, g->n[d]);
    }
    if (g->n[2] == 1 && !(g->gamma_hi == 1.0 && g->gamma_lo == 0.0))
        return set_error(ACX_EUNSUPPORTED, "2D grid (n[2] = 1): the shipped linear weights only (gamma = 1, 0)");
    if ((size_t)g->n[0] * g->n[1] > 0x7fffffffu) return set_error(ACX_EINVAL, "plane too large");
    if (g->system != ACX_WAVE2 && g->system != ACX_ACOUSTIC4)
        return set_error(ACX_EINVAL, "system %d unknown", g->system);
    if (!(g->h > 0)) return set_error(ACX_EINVAL, "h must be positive");
    for (int d = 0; d < 3; ++d) {
        if (g->system == ACX_ACOUSTIC4 && g->bc[d] == ACX_ADIABATIC)
            return set_error(ACX_EUNSUPPORTED, "ACOUSTIC4: axis %d: adiabatic walls are a WAVE2 rule; use reflective", d);
        if (g->system == ACX_WAVE2 && (g->bc[d] == ACX_REFLECTIVE || g->bc[d] == ACX_TRANSMISSIVE))
            return set_error(ACX_EUNSUPPORTED, "WAVE2: axis %d: reflective/transmissive walls are ACOUSTIC4 rules", d);
    }
    if (g->system == ACX_ACOUSTIC4) {
        if (!(g->rho0 > 0) || !(g->c0 > 0)) return set_error(ACX_EINVAL, "rho0 and c0 must be positive");
    }
    if (!(g->gamma_hi > 0 && g->gamma_hi <= 1) || !(g->gamma_lo >= 0 && g->gamma_lo <= 1))
        return set_error(ACX_EINVAL, "linear weights out of range");
    return ACX_OK;
}

static inline bool is_linear(const acx_grid *g)
{
    return g->gamma_hi == 1.0 && g->gamma_lo == 0.0 && !env_int("ACX_FORCE_CWENO", 0);
}

static inline int nvar_of(const acx_grid *g) { return g->system == ACX_ACOUSTIC4 ? 4 : 2; }

static void fill_lin(LinArgs &a, const acx_grid *g)
{
    memset(&a, 0, sizeof(a));
    a.nx = g->n[0]; a.ny = g->n[1]; a.nz = g->n[2];
    // z-chunk per workgroup: ACX_ZC_LINEAR, else 0 = chosen by the launcher
    // for the kernel's tile and occupancy (launch_linear)
    const int zc = env_int("ACX_ZC_LINEAR", 0);
    a.zc = zc < 1 ? 0 : (zc > g->n[2] ? g->n[2] : zc);
    for (int d = 0; d < 3; ++d) a.bc[d] = g->bc[d];
    a.K = g->rho0 * g->c0 * g->c0;
    a.invrho = 1.0 / g->rho0;
    a.s = g->c0;
    a.inv48h = 1.0 / (48.0 * g->h);
    a.lapc = g->c0 * g->c0 / (48.0 * g->h * g->h);
}

static void fill_cw(CwArgs &a, const acx_grid *g)
{
    memset(&a, 0, sizeof(a));
    a.nx = g->n[0]; a.ny = g->n[1]; a.nz = g->n[2];
    // z-chunk: ACX_ZC_CWENO, else 0 = chosen by the launcher (launch_cweno)
    const int zc = env_int("ACX_ZC_CWENO", 0);
    a.zc = zc < 1 ? 0 : (zc > g->n[2] ? g->n[2] : zc);
    for (int d = 0; d < 3; ++d) a.bc[d] = g->bc[d];
    a.K = g->rho0 * g->c0 * g->c0;
    a.invrho = 1.0 / g->rho0;
    a.s = g->c0;
    a.inv_h = 1.0 / g->h;
    a.c2 = g->c0 * g->c0 / (g->h * g->h);
    // linear weights, reconstruction.c:55-59
    a.g4 = g->gamma_hi;
    a.g3c = (1.0 - g->gamma_hi) * g->gamma_lo;
    a.g3s = (1. / 6.) * (1 - g->gamma_hi) * (1 - g->gamma_lo);
    a.inv_g4 = 1.0 / g->gamma_hi;
}

enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3 };

// One FP64 stage: OUT = a*U0 + b*Q + c*F(Q), epilogue chosen from (a, b).
static int stage64(const acx_grid *g, const double *Q, const double *U0, double *OUT,
                   double a, double b, double c, int epi, hipStream_t s)
{
    if (is_linear(g)) {
        LinArgs la;
        fill_lin(la, g);
        la.q = Q; la.u0 = U0; la.out = OUT;
        la.a = a; la.b = b; la.c = c;
        return launch_linear(la, g->system, epi, false, s);
    }
    CwArgs ca;
    fill_cw(ca, g);
    ca.q = Q; ca.u0 = U0; ca.out = OUT;
    ca.a = a; ca.b = b; ca.c = c;
    return launch_cweno(ca, g->system, epi, false, false, s);
}

static int stage32c(const acx_grid *g, const float *HI, const float *D, const float *LO, float *OUT,
                    float *OUT2, double b, double c, int final_stage, const uint8_t *flags, hipStream_t s)
{
    if (is_linear(g) && !flags) {
        LinArgs la;
        fill_lin(la, g);
        la.q = HI; la.d = D; la.lo = LO; la.out = OUT; la.out2 = OUT2;
        la.b = b; la.c = c; la.final_stage = final_stage;
        return launch_linear(la, g->system, EPI_F32C, true, s);
    }
    CwArgs ca;
    fill_cw(ca, g);
    ca.q = HI; ca.d = D; ca.lo = LO; ca.out = OUT; ca.out2 = OUT2;
    ca.b = b; ca.c = c; ca.final_stage = final_stage;
    ca.flags = flags;
    return launch_cweno(ca, g->system, EPI_F32C, true, flags != nullptr, s);
}

} // namespace acx

using namespace acx;

__global__ void k_point_add(double *x, double v) { *x += v; }

extern "C" {

const char *acx_last_error(void) { return g_err; }
int acx_version(void) { return 1; }

double acx_cfl_dt(const acx_grid *g, double cfl)
{
    // dt = CFL h / (D s_max), s_max = max(c0, c0) = c0 (LinearisedEuler/pde.c:66,87);
    // D = 3, or 2 on a 2D grid (2d/weno4/bubble/rhs_function.c:385)
    return (cfl * g->h) / ((g->n[2] == 1 ? 2.0 : 3.0) * g->c0);
}

int acx_rhs_f64(const acx_grid *g, const double *U, double *RHS, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!U || !RHS || U == RHS) return set_error(ACX_EINVAL, "U and RHS must be distinct device arrays");
    return stage64(g, U, nullptr, RHS, 0.0, 0.0, 1.0, EPI_RHS, (hipStream_t)stream);
}

int acx_stage_f64(const acx_grid *g, const double *Q, const double *U0, double *OUT,
                  double a, double b, double c, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!Q || !OUT || OUT == Q || (U0 && OUT == U0)) return set_error(ACX_EINVAL, "OUT must not alias Q or U0");
    int epi;
    if (a == 0.0 && b == 0.0 && c == 1.0) epi = EPI_RHS;
    else if (a == 0.0 && b == 1.0) epi = EPI_AXPY;
    else {
        if (a != 0.0 && !U0) return set_error(ACX_EINVAL, "U0 is NULL while a != 0");
        epi = EPI_COMB;
        if (!U0) U0 = Q;   // a == 0: the base is unused
    }
    return stage64(g, Q, U0, OUT, a, b, c, epi, (hipStream_t)stream);
}

int acx_ssp_rks3_step_f64(const acx_grid *g, double *U, double *W0, double *W1, double dt, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!U || !W0 || !W1 || U == W0 || U == W1 || W0 == W1) return set_error(ACX_EINVAL, "U, W0, W1 must be distinct");
    hipStream_t s = (hipStream_t)stream;
    // TSSSPRKS3, s = 4 (n = 2, r = 2): stage times 0, 1/2, 1, 1/2.
    if ((rc = stage64(g, U, nullptr, W0, 0.0, 1.0, dt / 2, EPI_AXPY, s))) return rc;
    if ((rc = stage64(g, W0, nullptr, W1, 0.0, 1.0, dt / 2, EPI_AXPY, s))) return rc;
    if ((rc = stage64(g, W1, U, W0, 1. * 2 / 3, (2 - 1.) * 1. / 3, (2 - 1.) * dt / (2 * 3), EPI_COMB, s))) return rc;
    return stage64(g, W0, nullptr, U, 0.0, 1.0, dt / 2, EPI_AXPY, s);
}

int acx_ssp_rks3_step_f32c(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                           float *LO_OUT, float *D0, float *D1, double dt, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!HI || !LO || !HI_OUT || !LO_OUT || !D0 || !D1 || HI == HI_OUT || LO == LO_OUT)
        return set_error(ACX_EINVAL, "f32c step: outputs must not alias the inputs");
    hipStream_t s = (hipStream_t)stream;
    // increments relative to the base state U = HI + LO:
    //   D0 = dt/2 F(U); D1 = D0 + dt/2 F(U+D0); D0 = D1/3 + dt/6 F(U+D1);
    //   U += D0 + dt/2 F(U+D0)   (two-sum into HI_OUT, LO_OUT)
    if ((rc = stage32c(g, HI, nullptr, LO, D0, nullptr, 0.0, dt / 2, 0, nullptr, s))) return rc;
    if ((rc = stage32c(g, HI, D0, LO, D1, nullptr, 1.0, dt / 2, 0, nullptr, s))) return rc;
    if ((rc = stage32c(g, HI, D1, LO, D0, nullptr, 1.0 / 3.0, dt / 6, 0, nullptr, s))) return rc;
    return stage32c(g, HI, D0, LO, HI_OUT, LO_OUT, 1.0, dt / 2, 1, nullptr, s);
}


static int point_index(const acx_grid *g, int i, int j, int k, int comp, size_t *idx)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (i < 0 || j < 0 || k < 0 || i >= g->n[0] || j >= g->n[1] || k >= g->n[2])
        return set_error(ACX_EINVAL, "cell (%d, %d, %d) outside the grid", i, j, k);
    if (comp < 0 || comp >= nvar_of(g)) return set_error(ACX_EINVAL, "component %d out of range", comp);
    *idx = (((size_t)k * g->n[1] + j) * g->n[0] + i) * nvar_of(g) + comp;
    return ACX_OK;
}

int acx_point_add_f64(const acx_grid *g, double *X, int i, int j, int k, int comp, double value, void *stream)
{
    size_t idx;
    int rc = point_index(g, i, j, k, comp, &idx);
    if (rc) return rc;
    if (!X) return set_error(ACX_EINVAL, "X is NULL");
    hipLaunchKernelGGL(k_point_add, dim3(1), dim3(1), 0, (hipStream_t)stream, X + idx, value);
    return hip_check(hipGetLastError(), "point add");
}

int acx_point_get_f64(const acx_grid *g, const double *X, int i, int j, int k, int comp, double *out, void *stream)
{
    size_t idx;
    int rc = point_index(g, i, j, k, comp, &idx);
    if (rc) return rc;
    if (!X || !out) return set_error(ACX_EINVAL, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    rc = hip_check(hipMemcpyAsync(out, X + idx, sizeof(double), hipMemcpyDeviceToHost, s), "point get");
    return rc ? rc : hip_check(hipStreamSynchronize(s), "point get");
}

} // extern "C"

namespace acx {

// Tile mask of the selective step: one byte per CWENO tile, kept per device
// and grown on demand (selective steps issued concurrently on one device
// share it and must not overlap).
static uint8_t *sel_mask_scratch(size_t n)
{
    static std::mutex mu;
    static uint8_t *buf[64] = {};
    static size_t cap[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (cap[dev] < n) {
        if (buf[dev]) (void)hipFree(buf[dev]);
        buf[dev] = nullptr;
        cap[dev] = 0;
        if (hipMalloc(&buf[dev], n) != hipSuccess) return nullptr;
        cap[dev] = n;
    }
    return buf[dev];
}

// Patch tiling: 30 x 14 x zc cells, zc read once per process, so a tile mask
// made by acx_sel_mask stays valid for every later selective step.
static void fill_cw_patch(CwArgs &a, const acx_grid *g)
{
    fill_cw(a, g);
    static const int zc = env_int("ACX_ZC_PATCH", 4);
    a.zc = zc < 1 ? 1 : (zc > g->n[2] ? g->n[2] : zc);
}

// Selective stage: the linear kernel writes every cell, then the CWENO kernel
// rewrites the tiles that hold a flagged cell or a face neighbour of one
// (same stream, so the second pass lands last).
static int stage32c_sel(const acx_grid *g, const float *HI, const float *D, const float *LO, float *OUT, float *OUT2,
                        double b, double c, int final_stage, const uint8_t *flags, const uint8_t *mask, hipStream_t s)
{
    LinArgs la;
    fill_lin(la, g);
    la.q = HI; la.d = D; la.lo = LO; la.out = OUT; la.out2 = OUT2;
    la.b = b; la.c = c; la.final_stage = final_stage;
    int rc = launch_linear(la, g->system, EPI_F32C, true, s);
    if (rc) return rc;
    CwArgs ca;
    fill_cw_patch(ca, g);
    ca.q = HI; ca.d = D; ca.lo = LO; ca.out = OUT; ca.out2 = OUT2;
    ca.b = b; ca.c = c; ca.final_stage = final_stage;
    ca.flags = flags;
    ca.tilemask = mask;
    return launch_cweno(ca, g->system, EPI_F32C, true, true, s);
}

} // namespace acx

// selective (learned-indicator) step, declared in acx_learn.h
static int check_sel(const acx_grid *g, const float *HI, const float *LO, const float *HI_OUT, const float *LO_OUT,
                     const float *D0, const float *D1, const uint8_t *flags)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!flags) return set_error(ACX_EINVAL, "flags is NULL");
    if (!HI || !LO || !HI_OUT || !LO_OUT || !D0 || !D1 || HI == HI_OUT || LO == LO_OUT)
        return set_error(ACX_EINVAL, "f32c step: outputs must not alias the inputs");
    if (g->system != ACX_ACOUSTIC4) return set_error(ACX_EUNSUPPORTED, "selective step: ACOUSTIC4 only");
    return ACX_OK;
}

// The four selective stages given the tile mask (null mask: the check mode
// ACX_SEL_FULL=1, the CWENO kernel over every tile and no linear pass).
static int sel_stages(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT, float *LO_OUT, float *D0,
                      float *D1, const uint8_t *flags, const uint8_t *mask, double dt, hipStream_t s)
{
    int rc;
    if (!mask) {
        if ((rc = stage32c(g, HI, nullptr, LO, D0, nullptr, 0.0, dt / 2, 0, flags, s))) return rc;
        if ((rc = stage32c(g, HI, D0, LO, D1, nullptr, 1.0, dt / 2, 0, flags, s))) return rc;
        if ((rc = stage32c(g, HI, D1, LO, D0, nullptr, 1.0 / 3.0, dt / 6, 0, flags, s))) return rc;
        return stage32c(g, HI, D0, LO, HI_OUT, LO_OUT, 1.0, dt / 2, 1, flags, s);
    }
    if ((rc = stage32c_sel(g, HI, nullptr, LO, D0, nullptr, 0.0, dt / 2, 0, flags, mask, s))) return rc;
    if ((rc = stage32c_sel(g, HI, D0, LO, D1, nullptr, 1.0, dt / 2, 0, flags, mask, s))) return rc;
    if ((rc = stage32c_sel(g, HI, D1, LO, D0, nullptr, 1.0 / 3.0, dt / 6, 0, flags, mask, s))) return rc;
    return stage32c_sel(g, HI, D0, LO, HI_OUT, LO_OUT, 1.0, dt / 2, 1, flags, mask, s);
}

extern "C" int acx_ssp_rks3_step_f32c_sel(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                                          float *LO_OUT, float *D0, float *D1, const uint8_t *flags,
                                          double dt, void *stream)
{
    int rc = check_sel(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (env_int("ACX_SEL_FULL", 0)) return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, nullptr, dt, s);
    CwArgs ca;
    fill_cw_patch(ca, g);
    uint8_t *mask = sel_mask_scratch(cweno_sel_tiles(ca));
    if (!mask) return set_error(ACX_EHIP, "selective step: tile mask allocation failed");
    if ((rc = launch_cweno_tilemask(ca, flags, mask, s))) return rc;
    return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, mask, dt, s);
}

extern "C" size_t acx_sel_mask_bytes(const acx_grid *g)
{
    if (check_grid(g)) return 0;
    CwArgs ca;
    fill_cw_patch(ca, g);
    return cweno_sel_tiles(ca);
}

extern "C" int acx_sel_mask(const acx_grid *g, const uint8_t *flags, uint8_t *mask, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!flags || !mask) return set_error(ACX_EINVAL, "selective mask: null pointer");
    CwArgs ca;
    fill_cw_patch(ca, g);
    return launch_cweno_tilemask(ca, flags, mask, (hipStream_t)stream);
}

extern "C" int acx_ssp_rks3_step_f32c_selm(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                                           float *LO_OUT, float *D0, float *D1, const uint8_t *flags,
                                           const uint8_t *mask, double dt, void *stream)
{
    int rc = check_sel(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags);
    if (rc) return rc;
    if (!mask) return set_error(ACX_EINVAL, "selective step: mask is NULL");
    hipStream_t s = (hipStream_t)stream;
    if (env_int("ACX_SEL_FULL", 0)) return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, nullptr, dt, s);
    return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, mask, dt, s);
}
