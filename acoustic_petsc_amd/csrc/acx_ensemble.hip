// acx_ensemble.hip -- persistent-kernel stepping of a batch of independent
// acoustic problems (include/acx_ensemble.h).
//
// Work items are (problem, column tile, z-chunk) triples numbered
// problem-major; a workgroup takes the next item from an atomic counter,
// runs the same tile code as the per-grid linear kernel (acx_linear_core.h),
// and asks again until the queue is empty, so every workgroup exits once the
// counter passes the item count.  The counter is cleared by a one-thread
// kernel before every stage (a kernel node, so the step can be captured in a
// hipGraph).
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "acx_linear_core.h"
#include "../../include/acx_ensemble.h"

namespace acx {

struct EnsProblem {
    size_t off;          // first cell
    int n[3];
    int ntx, nty, ntz;
    int first;           // first work item
};

constexpr int ENS_BX = 64;
constexpr int ENS_MAXP = 256;

struct EnsConfig { int by, zc; };
constexpr int ENS_NCONFIG = 8;
static const EnsConfig kConfigs[ENS_NCONFIG] = {{8, 16}, {8, 32}, {8, 64}, {16, 16}, {16, 32}, {16, 64},
                                                {8, 128}, {16, 128}};

// EPI: EPI_F32C for stages 1-3, EPI_F32CF for the stage that ends the step.
// Registers capped for 4 waves per SIMD: the final stage with the non-finite
// watch otherwise takes 133 VGPRs on the 8-row tiles, 3 waves per SIMD, and
// loses 7 % of the ensemble rate (measured on MI355X).
template <int BY, int EPI>
__global__ void __launch_bounds__(ENS_BX *BY) __attribute__((amdgpu_waves_per_eu(4)))
k_ensemble(const EnsProblem *__restrict__ probs, int nprob, int total, unsigned *__restrict__ counter, LinArgs base)
{
    using Tile = LinTile<float, ACX_ACOUSTIC4, ENS_BX, BY>;
    __shared__ float lds[2][Tile::NS][BY + 6][ENS_BX + 6];
    __shared__ int s_item;
    __shared__ int s_first[ENS_MAXP];   // first work item of every problem (ascending)
    const int tid = threadIdx.y * ENS_BX + threadIdx.x;
    for (int i = tid; i < nprob; i += ENS_BX * BY) s_first[i] = probs[i].first;
    for (;;) {
        if (tid == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads();   // every thread has read s_item before it is overwritten
        if (item >= total) break;
        // the problem of the item: the last p with s_first[p] <= item (binary search)
        int p = 0, hi = nprob - 1;
        while (p < hi) {
            const int mid = (p + hi + 1) >> 1;
            if (s_first[mid] <= item) p = mid;
            else hi = mid - 1;
        }
        const EnsProblem pr = probs[p];
        const int local = item - pr.first;
        LinArgs a = base;
        a.nx = pr.n[0];
        a.ny = pr.n[1];
        a.nz = pr.n[2];
        a.ntx = pr.ntx;
        a.nty = pr.nty;
        a.ntz = pr.ntz;
        const size_t o = pr.off * 4;
        a.q = static_cast<const float *>(base.q) + o;
        a.d = base.d ? static_cast<const float *>(base.d) + o : nullptr;
        a.lo = static_cast<const float *>(base.lo) + o;
        a.out = static_cast<float *>(base.out) + o;
        a.out2 = base.out2 ? static_cast<float *>(base.out2) + o : nullptr;
        const int x0 = (local % pr.ntx) * ENS_BX, y0 = ((local / pr.ntx) % pr.nty) * BY;
        const int zb = (local / (pr.ntx * pr.nty)) * a.zc;
        // loads one plane ahead, periodic, results streamed past L2
        linear_tile<float, ACX_ACOUSTIC4, ENS_BX, BY, EPI, 1, false, true>(a, x0, y0, zb, lds);
        __syncthreads();   // LDS is reused by the next item
    }
}

__global__ void k_clear(unsigned *c) { *c = 0u; }

} // namespace acx

using namespace acx;

struct acx_ensemble {
    int nprob;
    double h;
    std::vector<EnsProblem> probs;      // for the current configuration
    std::vector<int> dims;
    size_t ncell;
    EnsProblem *d_probs;
    unsigned *d_counter;
    float *hi[2], *lo[2], *d0, *d1;
    int cur;                           // index of the current (hi, lo)
    int config;
    int persistent_blocks[2];          // per rows-per-tile variant (8, 16)
    hipGraphExec_t graph;
    double graph_dt;
    int graph_config, graph_cur;
    hipStream_t graph_stream;
};

static int ens_layout(acx_ensemble *e)
{
    const EnsConfig c = kConfigs[e->config];
    int first = 0;
    size_t off = 0;
    e->probs.resize(e->nprob);
    for (int p = 0; p < e->nprob; ++p) {
        EnsProblem &q = e->probs[p];
        for (int d = 0; d < 3; ++d) q.n[d] = e->dims[3 * p + d];
        q.off = off;
        q.ntx = (q.n[0] + ENS_BX - 1) / ENS_BX;
        q.nty = (q.n[1] + c.by - 1) / c.by;
        const int zc = c.zc < q.n[2] ? c.zc : q.n[2];
        q.ntz = (q.n[2] + zc - 1) / zc;
        q.first = first;
        first += q.ntx * q.nty * q.ntz;
        off += (size_t)q.n[0] * q.n[1] * q.n[2];
    }
    return hip_check(hipMemcpy(e->d_probs, e->probs.data(), sizeof(EnsProblem) * e->nprob, hipMemcpyHostToDevice),
                     "ensemble table");
}

static int ens_total(const acx_ensemble *e)
{
    const EnsProblem &l = e->probs.back();
    return l.first + l.ntx * l.nty * l.ntz;
}

template <int BY>
static int ens_launch(acx_ensemble *e, const LinArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(1), 0, s, e->d_counter);
    const int nb = e->persistent_blocks[BY == 8 ? 0 : 1];
    if (a.final_stage)
        hipLaunchKernelGGL((k_ensemble<BY, EPI_F32CF>), dim3(nb), dim3(ENS_BX, BY), 0, s, e->d_probs, e->nprob,
                           ens_total(e), e->d_counter, a);
    else
        hipLaunchKernelGGL((k_ensemble<BY, EPI_F32C>), dim3(nb), dim3(ENS_BX, BY), 0, s, e->d_probs, e->nprob,
                           ens_total(e), e->d_counter, a);
    return hip_check(hipGetLastError(), "ensemble launch");
}

static int ens_stage(acx_ensemble *e, const float *D, float *OUT, float *OUT2, double b, double c, int final_stage,
                     hipStream_t s)
{
    LinArgs a;
    memset(&a, 0, sizeof(a));
    a.q = e->hi[e->cur];
    a.d = D;
    a.lo = e->lo[e->cur];
    a.out = OUT;
    a.out2 = OUT2;
    a.b = b;
    a.c = c;
    a.final_stage = final_stage;
    a.zc = kConfigs[e->config].zc;
    a.K = 1.0;
    a.invrho = 1.0;
    a.s = 1.0;
    a.inv48h = 1.0 / (48.0 * e->h);
    a.nf = final_stage ? nonfinite_word_current() : nullptr;   // the step's output is watched
    // bc = periodic (0) from memset
    return kConfigs[e->config].by == 8 ? ens_launch<8>(e, a, s) : ens_launch<16>(e, a, s);
}

static int ens_step(acx_ensemble *e, double dt, hipStream_t s)
{
    const int nxt = 1 - e->cur;
    int rc;
    // same increment sequence as acx_ssp_rks3_step_f32c
    if ((rc = ens_stage(e, nullptr, e->d0, nullptr, 0.0, dt / 2, 0, s))) return rc;
    if ((rc = ens_stage(e, e->d0, e->d1, nullptr, 1.0, dt / 2, 0, s))) return rc;
    if ((rc = ens_stage(e, e->d1, e->d0, nullptr, 1.0 / 3.0, dt / 6, 0, s))) return rc;
    if ((rc = ens_stage(e, e->d0, e->hi[nxt], e->lo[nxt], 1.0, dt / 2, 1, s))) return rc;
    e->cur = nxt;
    return ACX_OK;
}

extern "C" {

int acx_ensemble_num_configs(void) { return ENS_NCONFIG; }

int acx_ensemble_create(int nprob, const int32_t *n, double h, acx_ensemble **out)
{
    if (nprob < 1 || nprob > ENS_MAXP || !n || !out || !(h > 0)) return set_error(ACX_EINVAL, "ensemble: bad arguments");
    for (int i = 0; i < 3 * nprob; ++i)
        if (n[i] < 4) return set_error(ACX_EINVAL, "ensemble: at least 4 cells per axis");
    acx_ensemble *e = new acx_ensemble();
    e->nprob = nprob;
    e->h = h;
    e->dims.assign(n, n + 3 * nprob);
    e->ncell = 0;
    for (int p = 0; p < nprob; ++p) e->ncell += (size_t)n[3 * p] * n[3 * p + 1] * n[3 * p + 2];
    e->config = 1;   // rows 8, z-chunk 32 until tuned
    (void)nonfinite_word();   // the watch word exists before any capture
    e->graph = nullptr;
    e->graph_stream = nullptr;
    const size_t bytes = e->ncell * 4 * sizeof(float);
    int rc = ACX_OK;
    rc = rc ? rc : hip_check(hipMalloc(&e->d_probs, sizeof(EnsProblem) * nprob), "ensemble alloc");
    rc = rc ? rc : hip_check(hipMalloc(&e->d_counter, sizeof(unsigned)), "ensemble alloc");
    for (int k = 0; k < 2 && !rc; ++k) {
        rc = hip_check(hipMalloc(&e->hi[k], bytes), "ensemble alloc");
        rc = rc ? rc : hip_check(hipMalloc(&e->lo[k], bytes), "ensemble alloc");
    }
    rc = rc ? rc : hip_check(hipMalloc(&e->d0, bytes), "ensemble alloc");
    rc = rc ? rc : hip_check(hipMalloc(&e->d1, bytes), "ensemble alloc");
    if (!rc) {
        hipMemset(e->hi[0], 0, bytes);
        hipMemset(e->lo[0], 0, bytes);
        e->cur = 0;
        int dev = 0, cus = 0, b8 = 0, b16 = 0, f8 = 0, f16 = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        // persistent grid: blocks that fit at once, the fewer of the two epilogue kinds
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b8, k_ensemble<8, EPI_F32C>, ENS_BX * 8, 0);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b16, k_ensemble<16, EPI_F32C>, ENS_BX * 16, 0);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&f8, k_ensemble<8, EPI_F32CF>, ENS_BX * 8, 0);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&f16, k_ensemble<16, EPI_F32CF>, ENS_BX * 16, 0);
        b8 = f8 < b8 ? f8 : b8;
        b16 = f16 < b16 ? f16 : b16;
        e->persistent_blocks[0] = cus * (b8 > 0 ? b8 : 1);
        e->persistent_blocks[1] = cus * (b16 > 0 ? b16 : 1);
        rc = ens_layout(e);
    }
    if (rc) {
        acx_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return ACX_OK;
}

int acx_ensemble_destroy(acx_ensemble *e)
{
    if (!e) return ACX_OK;
    if (e->graph) hipGraphExecDestroy(e->graph);
    hipFree(e->d_probs);
    hipFree(e->d_counter);
    for (int k = 0; k < 2; ++k) {
        hipFree(e->hi[k]);
        hipFree(e->lo[k]);
    }
    hipFree(e->d0);
    hipFree(e->d1);
    delete e;
    return ACX_OK;
}

int acx_ensemble_offsets(const acx_ensemble *e, size_t *offsets)
{
    if (!e || !offsets) return set_error(ACX_EINVAL, "ensemble: bad arguments");
    for (int p = 0; p < e->nprob; ++p) offsets[p] = e->probs[p].off;
    offsets[e->nprob] = e->ncell;
    return ACX_OK;
}

int acx_ensemble_state(acx_ensemble *e, float **hi, float **lo)
{
    if (!e || !hi || !lo) return set_error(ACX_EINVAL, "ensemble: bad arguments");
    *hi = e->hi[e->cur];
    *lo = e->lo[e->cur];
    return ACX_OK;
}

int acx_ensemble_copy_in(acx_ensemble *e, const float *hi, const float *lo, void *stream)
{
    if (!e || !hi || !lo) return set_error(ACX_EINVAL, "ensemble: bad arguments");
    const size_t bytes = e->ncell * 4 * sizeof(float);
    hipStream_t s = (hipStream_t)stream;
    int rc = hip_check(hipMemcpyAsync(e->hi[e->cur], hi, bytes, hipMemcpyDeviceToDevice, s), "ensemble copy in");
    return rc ? rc : hip_check(hipMemcpyAsync(e->lo[e->cur], lo, bytes, hipMemcpyDeviceToDevice, s), "ensemble copy in");
}

int acx_ensemble_copy_out(const acx_ensemble *e, float *hi, float *lo, void *stream)
{
    if (!e || !hi || !lo) return set_error(ACX_EINVAL, "ensemble: bad arguments");
    const size_t bytes = e->ncell * 4 * sizeof(float);
    hipStream_t s = (hipStream_t)stream;
    int rc = hip_check(hipMemcpyAsync(hi, e->hi[e->cur], bytes, hipMemcpyDeviceToDevice, s), "ensemble copy out");
    return rc ? rc : hip_check(hipMemcpyAsync(lo, e->lo[e->cur], bytes, hipMemcpyDeviceToDevice, s), "ensemble copy out");
}

int acx_ensemble_set_config(acx_ensemble *e, int config)
{
    if (!e || config < 0 || config >= ENS_NCONFIG) return set_error(ACX_EINVAL, "ensemble: config out of range");
    e->config = config;
    return ens_layout(e);
}

int acx_ensemble_get_config(const acx_ensemble *e) { return e ? e->config : -1; }

int acx_ensemble_step(acx_ensemble *e, double dt, void *stream)
{
    if (!e) return set_error(ACX_EINVAL, "ensemble: null");
    if (nonfinite_raised()) return set_error(ACX_ENONFINITE, "ensemble step: a kernel on this device wrote a NaN or Inf");
    return ens_step(e, dt, (hipStream_t)stream);
}

int acx_ensemble_graph_steps(acx_ensemble *e, double dt, int nsteps, void *stream)
{
    if (!e || nsteps < 0 || nsteps % 2) return set_error(ACX_EINVAL, "ensemble: nsteps must be even");
    if (nonfinite_raised()) return set_error(ACX_ENONFINITE, "ensemble steps: a kernel on this device wrote a NaN or Inf");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (!e->graph || e->graph_dt != dt || e->graph_config != e->config || e->graph_cur != e->cur) {
        if (e->graph) {
            hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        hipStream_t cs;
        if ((rc = hip_check(hipStreamCreate(&cs), "ensemble capture stream"))) return rc;
        const int cur0 = e->cur;
        hipGraph_t g;
        rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
        if (!rc) rc = ens_step(e, dt, cs);
        if (!rc) rc = ens_step(e, dt, cs);
        int rc2 = hip_check(hipStreamEndCapture(cs, &g), "end capture");
        rc = rc ? rc : rc2;
        e->cur = cur0;   // capture did not run anything
        if (!rc) {
            rc = hip_check(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0), "graph instantiate");
            hipGraphDestroy(g);
        }
        hipStreamDestroy(cs);
        if (rc) return rc;
        e->graph_dt = dt;
        e->graph_config = e->config;
        e->graph_cur = e->cur;
    }
    for (int k = 0; k < nsteps; k += 2)
        if ((rc = hip_check(hipGraphLaunch(e->graph, s), "graph launch"))) return rc;
    // two steps per replay: the state is back in the same buffer
    return ACX_OK;
}

int acx_ensemble_autotune(acx_ensemble *e, double dt, void *stream, int *best, float *ms_per_step)
{
    if (!e) return set_error(ACX_EINVAL, "ensemble: null");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = e->ncell * 4 * sizeof(float);
    float *shi = nullptr, *slo = nullptr;
    int rc = hip_check(hipMalloc(&shi, bytes), "autotune alloc");
    if (!rc) rc = hip_check(hipMalloc(&slo, bytes), "autotune alloc");
    if (!rc) rc = hip_check(hipMemcpyAsync(shi, e->hi[e->cur], bytes, hipMemcpyDeviceToDevice, s), "autotune save");
    if (!rc) rc = hip_check(hipMemcpyAsync(slo, e->lo[e->cur], bytes, hipMemcpyDeviceToDevice, s), "autotune save");
    const int saved_cur = e->cur;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    float bestms = 1e30f;
    int bestc = e->config;
    for (int c = 0; c < ENS_NCONFIG && !rc; ++c) {
        if ((rc = acx_ensemble_set_config(e, c))) break;
        if ((rc = ens_step(e, dt, s))) break;   // warm-up
        hipEventRecord(a, s);
        for (int k = 0; k < 4 && !rc; ++k) rc = ens_step(e, dt, s);
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        ms /= 4.f;
        if (ms < bestms) {
            bestms = ms;
            bestc = c;
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (!rc) {
        // restore the state the tuning steps advanced
        e->cur = saved_cur;
        rc = hip_check(hipMemcpyAsync(e->hi[e->cur], shi, bytes, hipMemcpyDeviceToDevice, s), "autotune restore");
        if (!rc) rc = hip_check(hipMemcpyAsync(e->lo[e->cur], slo, bytes, hipMemcpyDeviceToDevice, s), "autotune restore");
        if (!rc) rc = hip_check(hipStreamSynchronize(s), "autotune sync");
        if (!rc) rc = acx_ensemble_set_config(e, bestc);
    }
    hipFree(shi);
    hipFree(slo);
    if (!rc) {
        if (best) *best = bestc;
        if (ms_per_step) *ms_per_step = bestms;
    }
    return rc;
}

} // extern "C"
