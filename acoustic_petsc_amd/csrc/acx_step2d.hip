// acx_step2d.hip -- one whole SSP-RK3 (s = 4) step of the linear ACOUSTIC4
// operator on a small periodic one-plane (2D) grid, in one launch.
//
// On a 2D grid of 256^2 cells a stage launch holds 64 K cells: the four
// stage launches of a step are bound by launch and ramp latency, not by
// memory (about 5.5 us each).  Here a block takes a 16 x 16 tile of the
// output and the state on the tile with a 12-cell halo (40 x 40 cells) into
// LDS, then runs the four stages in LDS, each on a region 6 cells narrower
// than the last (34^2, 28^2, 22^2, 16^2): every stage reads a 3-cell halo
// of the one before (the 7-point A7/B7 stencils of the linear CWENO4
// reconstruction + LLF flux, 3d/weno3/LinearisedEuler/rhs_function.c with
// the shipped linear weights; on one plane every z term is exactly zero).
// The cells of the halo rings are computed by several blocks (2.6 x the
// stage evaluations of the tile), which costs LDS and VALU time, not HBM.
// Stage coefficients as in acx_ssp_rks3_step_f64 (TSSSPRKS3, 4 stages):
// W0 = U + dt/2 F(U); W1 = W0 + dt/2 F(W0); W0 = 2/3 U + 1/3 W1 + dt/6 F(W1);
// U' = W0 + dt/2 F(W0).  The step reads IN and writes OUT (distinct arrays:
// neighbouring blocks read IN on their halo while this block writes).
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

struct Step2d {
    static constexpr int T = 16;                   // output tile edge
    static constexpr int H = 12;                   // halo: 4 stages x 3 cells
    static constexpr int R0 = T + 2 * H;           // 40: state region
    static constexpr int R1 = R0 - 6, R2 = R1 - 6, R3 = R2 - 6;   // 34, 28, 22 (stage outputs 1-3)
    static constexpr int NT = 1024;                // threads per block
    // LDS: buffer A holds the state (R0^2 cells), later stage 2 (R2^2);
    // buffer B stage 1 (R1^2), later stage 3 (R3^2).  Each as two arrays of
    // 16-byte halves, (p, u) and (v, w), so that a wavefront reading
    // consecutive cells reads consecutive 16-byte words.
    static constexpr int NA = R0 * R0, NB = R1 * R1;
};

// RHS of cell (r, c) of a region of width w held as halves lo[], hi[] whose
// cell (r + 3, c + 3) is the centre; the LinArgs coefficients as in the
// stage kernels.
__device__ __forceinline__ void rhs2d(const double2 *__restrict__ lo, const double2 *__restrict__ hi, int w, int r,
                                      int c, const LinArgs &p, double (&out)[4], double (&q)[4])
{
    double2 xa[7], xb[7], ya[7], yb[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int ix = (r + 3) * w + (c + k), iy = (r + k) * w + (c + 3);
        xa[k] = lo[ix];
        xb[k] = hi[ix];
        ya[k] = lo[iy];
        yb[k] = hi[iy];
    }
    double X[4][7], Y[4][7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        X[0][k] = xa[k].x; X[1][k] = xa[k].y; X[2][k] = xb[k].x; X[3][k] = xb[k].y;
        Y[0][k] = ya[k].x; Y[1][k] = ya[k].y; Y[2][k] = yb[k].x; Y[3][k] = yb[k].y;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) q[v] = X[v][3];
    double accB[4], dAp0, dAp1, dAvel;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        accB[v] = stencil_B7(X[v][0], X[v][1], X[v][2], q[v], X[v][4], X[v][5], X[v][6]) +
                  stencil_B7(Y[v][0], Y[v][1], Y[v][2], q[v], Y[v][4], Y[v][5], Y[v][6]);
    dAp0 = stencil_A7(X[0][0], X[0][1], X[0][2], X[0][4], X[0][5], X[0][6]);
    dAp1 = stencil_A7(Y[0][0], Y[0][1], Y[0][2], Y[0][4], Y[0][5], Y[0][6]);
    dAvel = stencil_A7(X[1][0], X[1][1], X[1][2], X[1][4], X[1][5], X[1][6]) +
            stencil_A7(Y[2][0], Y[2][1], Y[2][2], Y[2][4], Y[2][5], Y[2][6]);
    const double inv48h = p.inv48h, Kc = p.K, irho = p.invrho, sLLF = p.s;
    out[0] = inv48h * (sLLF * accB[0] - Kc * dAvel);
    out[1] = inv48h * (sLLF * accB[1] - irho * dAp0);
    out[2] = inv48h * (sLLF * accB[2] - irho * dAp1);
    out[3] = inv48h * (sLLF * accB[3]);
}

__global__ void __launch_bounds__(Step2d::NT) k_step2d(const double *__restrict__ IN, double *__restrict__ OUT,
                                                      LinArgs p, double dt)
{
    using S = Step2d;
    __shared__ double2 a_lo[S::NA], a_hi[S::NA];
    __shared__ double2 b_lo[S::NB], b_hi[S::NB];
    const int nx = p.nx, ny = p.ny;
    const int ntx = (nx + S::T - 1) / S::T;
    const int x0 = (blockIdx.x % ntx) * S::T, y0 = (blockIdx.x / ntx) * S::T;
    const int tid = threadIdx.x;
    auto M = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };

    // the state on the tile and its 12-cell halo (periodic wrap)
    for (int e = tid; e < S::NA; e += S::NT) {
        const int r = e / S::R0, c = e - r * S::R0;
        const size_t g = ((size_t)M(y0 - S::H + r, ny) * nx + M(x0 - S::H + c, nx)) * 4;
        a_lo[e] = *reinterpret_cast<const double2 *>(IN + g);
        a_hi[e] = *reinterpret_cast<const double2 *>(IN + g + 2);
    }
    __syncthreads();
    // the state on the stage-3 region (22^2, offset 9 in the 40^2 region),
    // kept in registers for the stage-3 combination (at most one cell per thread)
    double u0[4] = {0.0, 0.0, 0.0, 0.0};
    if (tid < S::R3 * S::R3) {
        const int r = tid / S::R3, c = tid - r * S::R3, e = (r + 9) * S::R0 + (c + 9);
        u0[0] = a_lo[e].x; u0[1] = a_lo[e].y; u0[2] = a_hi[e].x; u0[3] = a_hi[e].y;
    }
    const double h2 = dt / 2;
    // stage 1: A (40^2) -> B (34^2)
    for (int e = tid; e < S::R1 * S::R1; e += S::NT) {
        const int r = e / S::R1, c = e - r * S::R1;
        double f[4], q[4];
        rhs2d(a_lo, a_hi, S::R0, r, c, p, f, q);
        b_lo[e] = make_double2(q[0] + h2 * f[0], q[1] + h2 * f[1]);
        b_hi[e] = make_double2(q[2] + h2 * f[2], q[3] + h2 * f[3]);
    }
    __syncthreads();
    // stage 2: B (34^2) -> A (28^2)
    for (int e = tid; e < S::R2 * S::R2; e += S::NT) {
        const int r = e / S::R2, c = e - r * S::R2;
        double f[4], q[4];
        rhs2d(b_lo, b_hi, S::R1, r, c, p, f, q);
        a_lo[e] = make_double2(q[0] + h2 * f[0], q[1] + h2 * f[1]);
        a_hi[e] = make_double2(q[2] + h2 * f[2], q[3] + h2 * f[3]);
    }
    __syncthreads();
    // stage 3: A (28^2) + U -> B (22^2); one cell per thread
    const double ca = 1. * 2 / 3, cb = (2 - 1.) * 1. / 3, cc = (2 - 1.) * dt / (2 * 3);
    if (tid < S::R3 * S::R3) {
        const int r = tid / S::R3, c = tid - r * S::R3;
        double f[4], q[4], o[4];
        rhs2d(a_lo, a_hi, S::R2, r, c, p, f, q);
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = ca * u0[v] + cb * q[v] + cc * f[v];
        b_lo[tid] = make_double2(o[0], o[1]);
        b_hi[tid] = make_double2(o[2], o[3]);
    }
    __syncthreads();
    // stage 4: B (22^2) -> OUT (16^2)
    bool nfbad = false;
    if (tid < S::T * S::T) {
        const int r = tid / S::T, c = tid - r * S::T;
        double f[4], q[4];
        rhs2d(b_lo, b_hi, S::R3, r, c, p, f, q);
        Cell<double, 4> o;
#pragma unroll
        for (int v = 0; v < 4; ++v) o.v[v] = q[v] + h2 * f[v];
        const int gx = x0 + c, gy = y0 + r;
        if (gx < nx && gy < ny) {
            if (p.nf) finite_acc<double, 4>(nfbad, o);
            store_cell<double, 4>(OUT, (size_t)gy * nx + gx, o);
        }
    }
    watch_flush(p.nf, nfbad);
}

int launch_step2d(const LinArgs &a, const double *IN, double *OUT, double dt, hipStream_t s)
{
    const int ntx = (a.nx + Step2d::T - 1) / Step2d::T, nty = (a.ny + Step2d::T - 1) / Step2d::T;
    hipLaunchKernelGGL(k_step2d, dim3((unsigned)(ntx * nty)), dim3(Step2d::NT), 0, s, IN, OUT, a, dt);
    return hip_check(hipGetLastError(), "k_step2d launch");
}

} // namespace acx
