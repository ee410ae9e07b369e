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
#include "../../include/acx_learn.h"

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

// Non-finite watch: one pinned, device-mapped host word per device.  Stage
// kernels store 1 into it when they write a NaN or +-Inf (watch_finite in
// acx_device.h); the host reads it without synchronising.
struct NfWord {
    volatile unsigned *host;
    unsigned *dev;
};
static NfWord g_nf[64];
static std::mutex g_nf_mu;

static NfWord *nf_for_current(bool create)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    NfWord &w = g_nf[dev];
    if (w.dev || !create) return w.dev ? &w : nullptr;
    std::lock_guard<std::mutex> lk(g_nf_mu);
    if (!w.dev) {
        void *h = nullptr, *d = nullptr;
        if (hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return nullptr;
        if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) return nullptr;
        *(volatile unsigned *)h = 0u;
        w.host = (volatile unsigned *)h;
        w.dev = (unsigned *)d;
    }
    return &w;
}

unsigned *nonfinite_word()
{
    NfWord *w = nf_for_current(true);
    return w ? w->dev : nullptr;
}

unsigned *nonfinite_word_current()
{
    NfWord *w = nf_for_current(false);
    return w ? w->dev : nullptr;
}

bool nonfinite_raised()
{
    NfWord *w = nf_for_current(false);
    return w && *w->host != 0u;
}

// Entry check of every RHS / stage / step call (creates the device's watch
// word on its first call there, so the kernels it launches carry it).
static int nonfinite_entry(const char *what)
{
    (void)nonfinite_word();
    if (nonfinite_raised())
        return set_error(ACX_ENONFINITE, "%s: a kernel on this device wrote a NaN or Inf "
                                         "(acx_nonfinite_clear resets the watch)", what);
    return ACX_OK;
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
    a.nf = nonfinite_word_current();
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
    if (g->n[2] == 1) {
        // One plane (2D grid): the 2D CWENO4 of 2d/weno4/bubble/reconstruction.c:128-251
        // has five quadratic stencils (central, +-x, +-y) with weights
        // (1 - gamma_hi) gamma_lo and (1 - gamma_hi)(1 - gamma_lo)/4 (:130-140).
        // On a z-invariant plane the 3D stencils +-z equal the central one
        // and every z term of the indicators and of R4 vanishes (the 2D IS_R4
        // of :154-156 is the 3D one without them), so the 3D scheme is the 2D
        // scheme when the sided weight is the 2D one and the central weight
        // gives up the share of the two z stencils: the three have the same
        // indicator on one plane, so only the sum of their linear weights (the
        // 2D central weight) enters the result, whatever its split.
        a.g3s = 0.25 * (1 - g->gamma_hi) * (1 - g->gamma_lo);
        a.g3c = (1.0 - g->gamma_hi) * g->gamma_lo - 2.0 * a.g3s;
    }
    a.inv_g4 = 1.0 / g->gamma_hi;
    a.nf = nonfinite_word_current();
}

enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3 };

// One FP64 stage: OUT = a*U0 + b*Q + c*F(Q), epilogue chosen from (a, b).
// watch: the stage's outputs feed the non-finite watch (a step watches only
// its last stage: a NaN or Inf of an earlier stage reaches its output).
// Small periodic one-plane ACOUSTIC4 grids with the shipped weights: a step
// is one launch (k_step2d; up to 2^20 cells, where the four stage launches
// are bound by launch latency rather than memory).
static bool whole_step_2d(const acx_grid *g)
{
    return g->n[2] == 1 && g->system == ACX_ACOUSTIC4 && is_linear(g) && g->bc[0] == ACX_PERIODIC &&
           g->bc[1] == ACX_PERIODIC && g->bc[2] == ACX_PERIODIC && (size_t)g->n[0] * g->n[1] <= (1u << 20);
}

static int step2d(const acx_grid *g, const double *IN, double *OUT, double dt, hipStream_t s)
{
    LinArgs la;
    fill_lin(la, g);
    return launch_step2d(la, IN, OUT, dt, s);
}

static int stage64(const acx_grid *g, const double *Q, const double *U0, double *OUT,
                   double a, double b, double c, int epi, hipStream_t s, bool watch = true)
{
    if (is_linear(g)) {
        LinArgs la;
        fill_lin(la, g);
        la.q = Q; la.u0 = U0; la.out = OUT;
        la.a = a; la.b = b; la.c = c;
        if (!watch) la.nf = nullptr;
        return launch_linear(la, g->system, epi, false, s);
    }
    CwArgs ca;
    fill_cw(ca, g);
    ca.q = Q; ca.u0 = U0; ca.out = OUT;
    ca.a = a; ca.b = b; ca.c = c;
    if (!watch) ca.nf = nullptr;
    return launch_cweno(ca, g->system, epi, false, false, s);
}

// One compensated FP32 stage; only the stage that ends the step is watched.
static int stage32c(const acx_grid *g, const float *HI, const float *D, const float *LO, float *OUT,
                    float *OUT2, double b, double c, int final_stage, const uint8_t *flags, hipStream_t s)
{
    if (is_linear(g) && !flags) {
        LinArgs la;
        fill_lin(la, g);
        la.q = HI; la.d = D; la.lo = LO; la.out = OUT; la.out2 = OUT2;
        la.b = b; la.c = c; la.final_stage = final_stage;
        if (!final_stage) la.nf = nullptr;
        return launch_linear(la, g->system, EPI_F32C, true, s);
    }
    CwArgs ca;
    fill_cw(ca, g);
    ca.q = HI; ca.d = D; ca.lo = LO; ca.out = OUT; ca.out2 = OUT2;
    ca.b = b; ca.c = c; ca.final_stage = final_stage;
    ca.flags = flags;
    if (!final_stage) ca.nf = nullptr;
    return launch_cweno(ca, g->system, EPI_F32C, true, flags != nullptr, s);
}

// The six arrays of a compensated FP32 step: all present and pairwise
// distinct (the 2.5D march reads HI and the increments of neighbouring tiles
// as halo while other tiles write the outputs).
static int check_f32c_arrays(const float *HI, const float *LO, const float *HI_OUT, const float *LO_OUT,
                             const float *D0, const float *D1)
{
    const float *p[6] = {HI, LO, HI_OUT, LO_OUT, D0, D1};
    for (int i = 0; i < 6; ++i) {
        if (!p[i]) return set_error(ACX_EINVAL, "f32c step: null array");
        for (int j = 0; j < i; ++j)
            if (p[i] == p[j]) return set_error(ACX_EINVAL, "f32c step: HI, LO, HI_OUT, LO_OUT, D0, D1 must be distinct");
    }
    return ACX_OK;
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
    if ((rc = nonfinite_entry("rhs"))) return rc;
    return stage64(g, U, nullptr, RHS, 0.0, 0.0, 1.0, EPI_RHS, (hipStream_t)stream);
}

int acx_stage_f64(const acx_grid *g, const double *Q, const double *U0, double *OUT,
                  double a, double b, double c, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!Q || !OUT || OUT == Q || (U0 && OUT == U0)) return set_error(ACX_EINVAL, "OUT must not alias Q or U0");
    if ((rc = nonfinite_entry("stage"))) return rc;
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
    if ((rc = nonfinite_entry("ssp_rks3_step_f64"))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (whole_step_2d(g)) {
        // the step in one launch into W0, copied back to U
        if ((rc = step2d(g, U, W0, dt, s))) return rc;
        return acx_vec_copy_f64(W0, U, (size_t)g->n[0] * g->n[1] * 4, s);
    }
    // TSSSPRKS3, s = 4 (n = 2, r = 2): stage times 0, 1/2, 1, 1/2.
    if ((rc = stage64(g, U, nullptr, W0, 0.0, 1.0, dt / 2, EPI_AXPY, s, false))) return rc;
    if ((rc = stage64(g, W0, nullptr, W1, 0.0, 1.0, dt / 2, EPI_AXPY, s, false))) return rc;
    if ((rc = stage64(g, W1, U, W0, 1. * 2 / 3, (2 - 1.) * 1. / 3, (2 - 1.) * dt / (2 * 3), EPI_COMB, s, false)))
        return rc;
    return stage64(g, W0, nullptr, U, 0.0, 1.0, dt / 2, EPI_AXPY, s);
}

} // extern "C"

namespace acx {

// The stage launches of `steps` FP64 steps, captured once per (device,
// grid, arrays, dt, steps) and replayed; one cached graph per device.
constexpr int kMaxGraphSteps = 32;
struct StepGraph {
    bool valid = false;
    acx_grid g;
    double *U = nullptr, *W0 = nullptr, *W1 = nullptr;
    double dt = 0.0;
    int steps = 0;
    hipGraphExec_t exec = nullptr;
};
static StepGraph g_step_graph[64];
static std::mutex g_step_graph_mu;

static bool same_grid(const acx_grid &a, const acx_grid &b) { return memcmp(&a, &b, sizeof(acx_grid)) == 0; }

static int step_graph_for(const acx_grid *g, double *U, double *W0, double *W1, double dt, int steps,
                          hipGraphExec_t *out)
{
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "step graph: device");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return set_error(ACX_EUNSUPPORTED, "step graph: device %d", dev);
    StepGraph &sg = g_step_graph[dev];
    if (sg.valid && same_grid(sg.g, *g) && sg.U == U && sg.W0 == W0 && sg.W1 == W1 && sg.dt == dt &&
        sg.steps == steps) {
        *out = sg.exec;
        return ACX_OK;
    }
    if (sg.exec) {
        (void)hipGraphExecDestroy(sg.exec);
        sg.exec = nullptr;
    }
    sg.valid = false;
    hipStream_t cs;
    if ((rc = hip_check(hipStreamCreate(&cs), "step graph: capture stream"))) return rc;
    hipGraph_t graph = nullptr;
    rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed), "step graph: begin capture");
    if (!rc) {
        // the stage sequence of acx_ssp_rks3_step_f64, `steps` times (a
        // one-launch step alternates between U and W0, and ends in U)
        const bool whole = whole_step_2d(g);
        for (int k = 0; k < steps && !rc && whole; ++k)
            rc = step2d(g, k % 2 ? W0 : U, k % 2 ? U : W0, dt, cs);
        if (!rc && whole && steps % 2) rc = acx_vec_copy_f64(W0, U, (size_t)g->n[0] * g->n[1] * 4, cs);
        for (int k = 0; k < steps && !rc && !whole; ++k) {
            rc = stage64(g, U, nullptr, W0, 0.0, 1.0, dt / 2, EPI_AXPY, cs, false);
            if (!rc) rc = stage64(g, W0, nullptr, W1, 0.0, 1.0, dt / 2, EPI_AXPY, cs, false);
            if (!rc)
                rc = stage64(g, W1, U, W0, 1. * 2 / 3, (2 - 1.) * 1. / 3, (2 - 1.) * dt / (2 * 3), EPI_COMB, cs,
                             false);
            if (!rc) rc = stage64(g, W0, nullptr, U, 0.0, 1.0, dt / 2, EPI_AXPY, cs);
        }
        const int re = hip_check(hipStreamEndCapture(cs, &graph), "step graph: end capture");
        if (!rc) rc = re;
    }
    if (!rc) rc = hip_check(hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0), "step graph: instantiate");
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    sg.g = *g;
    sg.U = U;
    sg.W0 = W0;
    sg.W1 = W1;
    sg.dt = dt;
    sg.steps = steps;
    sg.valid = true;
    *out = sg.exec;
    return ACX_OK;
}

} // namespace acx

extern "C" {

int acx_ssp_rks3_steps_graph_f64(const acx_grid *g, double *U, double *W0, double *W1, double dt, int nsteps,
                                 void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!U || !W0 || !W1 || U == W0 || U == W1 || W0 == W1) return set_error(ACX_EINVAL, "U, W0, W1 must be distinct");
    if (nsteps < 0) return set_error(ACX_EINVAL, "nsteps must be >= 0");
    if ((rc = nonfinite_entry("ssp_rks3_steps_graph_f64"))) return rc;   // creates the watch word before capture
    if (nsteps == 0) return ACX_OK;
    std::lock_guard<std::mutex> lk(g_step_graph_mu);
    // one graph of up to kMaxGraphSteps steps, replayed; the remainder from
    // a graph of its own size (both are cached in turn)
    const int per = nsteps < kMaxGraphSteps ? nsteps : kMaxGraphSteps;
    hipGraphExec_t exec = nullptr;
    if ((rc = step_graph_for(g, U, W0, W1, dt, per, &exec))) return rc;
    for (int k = 0; k + per <= nsteps; k += per)
        if ((rc = hip_check(hipGraphLaunch(exec, (hipStream_t)stream), "step graph: launch"))) return rc;
    const int rem = nsteps % per;
    if (rem) {
        if ((rc = step_graph_for(g, U, W0, W1, dt, rem, &exec))) return rc;
        rc = hip_check(hipGraphLaunch(exec, (hipStream_t)stream), "step graph: launch");
    }
    return rc;
}

int acx_ssp_rks3_step_f32c(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                           float *LO_OUT, float *D0, float *D1, double dt, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_f32c_arrays(HI, LO, HI_OUT, LO_OUT, D0, D1))) return rc;
    if ((rc = nonfinite_entry("ssp_rks3_step_f32c"))) return rc;
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

int acx_nonfinite_check(void *stream)
{
    int rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "nonfinite check");
    if (rc) return rc;
    return nonfinite_entry("nonfinite check");
}

int acx_nonfinite_clear(void)
{
    NfWord *w = nf_for_current(true);
    if (!w) return set_error(ACX_EHIP, "nonfinite clear: watch word unavailable");
    *w->host = 0u;
    return ACX_OK;
}

} // extern "C"

namespace acx {

// Patch tiling: 30 x 14 x zc cells, zc read once per process, so a tile mask
// made by acx_sel_mask stays valid for every later selective step.  Measured
// on a steep pulse (flags of 1.2-1.9 % of the (cell, component) pairs) at
// 256^3 and 384^3: zc = 8 leaves about one round of resident patch blocks
// (508 tiles at 384^3) and took the step 3.32 -> 3.22 ms at 384^3 against
// zc = 4 (968 tiles, two rounds); zc = 2 and 16 were slower (3.52, 3.48 ms).
static void fill_cw_patch(CwArgs &a, const acx_grid *g)
{
    fill_cw(a, g);
    static const int zc = env_int("ACX_ZC_PATCH", 8);
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
    if (!final_stage) la.nf = nullptr;
    int rc = launch_linear(la, g->system, EPI_F32C, true, s);
    if (rc) return rc;
    CwArgs ca;
    fill_cw_patch(ca, g);
    ca.q = HI; ca.d = D; ca.lo = LO; ca.out = OUT; ca.out2 = OUT2;
    ca.b = b; ca.c = c; ca.final_stage = final_stage;
    ca.flags = flags;
    ca.tlist = reinterpret_cast<const int *>(mask + sel_list_offset(cweno_sel_tiles(ca)));
    if (!final_stage) ca.nf = nullptr;
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
    if ((rc = check_f32c_arrays(HI, LO, HI_OUT, LO_OUT, D0, D1))) return rc;
    if (g->system != ACX_ACOUSTIC4) return set_error(ACX_EUNSUPPORTED, "selective step: ACOUSTIC4 only");
    return nonfinite_entry("selective step");
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
    // the tile mask of this call, from the stream-ordered pool of s
    uint8_t *mask = nullptr;
    if ((rc = hip_check(hipMallocAsync((void **)&mask, sel_buffer_bytes(cweno_sel_tiles(ca)), s),
                        "selective step: tile mask")))
        return rc;
    rc = launch_cweno_tilemask(ca, flags, mask, s);
    if (!rc) rc = sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, mask, dt, s);
    const int rf = hip_check(hipFreeAsync(mask, s), "selective step: tile mask free");
    return rc ? rc : rf;
}

extern "C" size_t acx_sel_mask_bytes(const acx_grid *g)
{
    if (check_grid(g)) return 0;
    CwArgs ca;
    fill_cw_patch(ca, g);
    return sel_buffer_bytes(cweno_sel_tiles(ca));
}

extern "C" int acx_sel_mask(const acx_grid *g, const uint8_t *flags, uint8_t *mask, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!flags || !mask) return set_error(ACX_EINVAL, "selective mask: null pointer");
    if ((uintptr_t)mask % 4) return set_error(ACX_EINVAL, "selective mask: buffer not 4-byte aligned");
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
    if ((uintptr_t)mask % 4) return set_error(ACX_EINVAL, "selective step: mask not 4-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    if (env_int("ACX_SEL_FULL", 0)) return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, nullptr, dt, s);
    return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, mask, dt, s);
}

// ---- selective step without allocation -------------------------------------
//
// The tile mask lives in a file-scope device array of fixed size; each device
// holds its own copy.  Host code never passes the array's name to a kernel
// (on the host it is only a stand-in symbol): the device address is looked up
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) uint8_t scratch[acx::sel_buffer_bytes(ACX_SEL_NOALLOC_MAX_TILES)];
static uint8_t *g_scratch_ptr[64];
static std::mutex g_scratch_mu;

static int scratch_for_current(uint8_t **out)
{
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "selective step (no allocation): device");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return set_error(ACX_EUNSUPPORTED, "selective step (no allocation): device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if ((rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "selective step (no allocation): scratch address")))
            return rc;
        g_scratch_ptr[dev] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[dev];
    return ACX_OK;
}

extern "C" int acx_ssp_rks3_step_f32c_sel_noalloc(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                                                  float *LO_OUT, float *D0, float *D1, const uint8_t *flags,
                                                  double dt, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!flags) return set_error(ACX_EINVAL, "flags is NULL");
    if ((rc = check_f32c_arrays(HI, LO, HI_OUT, LO_OUT, D0, D1))) return rc;
    if (g->system != ACX_ACOUSTIC4) return set_error(ACX_EUNSUPPORTED, "selective step: ACOUSTIC4 only");
    CwArgs ca;
    fill_cw_patch(ca, g);
    const size_t tiles = cweno_sel_tiles(ca);
    if (tiles > ACX_SEL_NOALLOC_MAX_TILES)
        return set_error(ACX_EINVAL, "selective step (no allocation): %zu patch tiles, at most %d fit the static scratch",
                         tiles, ACX_SEL_NOALLOC_MAX_TILES);
    // the watch word is not created here (that would allocate): it is used
    // when another entry point has created it on this device
    if (nonfinite_raised())
        return set_error(ACX_ENONFINITE, "selective step: a kernel on this device wrote a NaN or Inf");
    uint8_t *mask = nullptr;
    if ((rc = scratch_for_current(&mask))) return rc;
    hipStream_t s = (hipStream_t)stream;
    // every mask entry of the grid's tiles is written by the marking kernel
    if ((rc = launch_cweno_tilemask(ca, flags, mask, s))) return rc;
    return sel_stages(g, HI, LO, HI_OUT, LO_OUT, D0, D1, flags, mask, dt, s);
}
