// acx_misc.hip -- FP64 reductions, the CTCS leapfrog sibling and host-side
// synthetic fields.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

// ---- deterministic two-pass FP64 reductions ------------------------------
// pass 1: each block reduces a grid-stride slice to (sum, max) with wavefront
// shuffles; pass 2: one block reduces the partials in a fixed order.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

// max that keeps a NaN (fmax drops it): a NaN anywhere makes the result NaN,
// as PETSc's VecNorm(NORM_INFINITY) reports it
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ __forceinline__ void wave_reduce(double &s, double &m)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        m = nmax(m, __shfl_down(m, off, 64));
    }
}

__device__ __forceinline__ void block_reduce(double &s, double &m)
{
    __shared__ double ss[RED_THREADS / 64], sm[RED_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    wave_reduce(s, m);
    if (lane == 0) { ss[wv] = s; sm[wv] = m; }
    __syncthreads();
    if (wv == 0) {
        s = lane < RED_THREADS / 64 ? ss[lane] : 0.0;
        m = lane < RED_THREADS / 64 ? sm[lane] : 0.0;
        wave_reduce(s, m);
    }
}

// ---- vector copy ------------------------------------------------------------
// Y = X with nontemporal 16-byte loads and stores, one pair per thread: a
// block covers CP_THREADS * CP_UNROLL pairs.  Measured on MI355X (read +
// write rate, three alternations on one box): a 256^3 FP64 ACOUSTIC4 state
// 6.29 -> 6.52 TB/s and a 384^3 one 6.10 -> 6.43 TB/s against four pairs per
// thread (each thread's loads issued before its first store).  k_copy1 is
// the scalar form for pointers off 16-byte alignment and for the odd last
// element.
constexpr int CP_THREADS = 256, CP_UNROLL = 1;

__global__ void __launch_bounds__(CP_THREADS) k_copy2(const acx_dbl2 *__restrict__ X, acx_dbl2 *__restrict__ Y, size_t n2)
{
    const size_t base = (size_t)blockIdx.x * (CP_THREADS * CP_UNROLL) + threadIdx.x;
    acx_dbl2 v[CP_UNROLL];
#pragma unroll
    for (int u = 0; u < CP_UNROLL; ++u) {
        const size_t i = base + (size_t)u * CP_THREADS;
        if (i < n2) v[u] = __builtin_nontemporal_load(X + i);
    }
#pragma unroll
    for (int u = 0; u < CP_UNROLL; ++u) {
        const size_t i = base + (size_t)u * CP_THREADS;
        if (i < n2) __builtin_nontemporal_store(v[u], Y + i);
    }
}

__global__ void __launch_bounds__(CP_THREADS) k_copy1(const double *__restrict__ X, double *__restrict__ Y, size_t n)
{
    const size_t base = (size_t)blockIdx.x * (CP_THREADS * CP_UNROLL) + threadIdx.x;
    double v[CP_UNROLL];
#pragma unroll
    for (int u = 0; u < CP_UNROLL; ++u) {
        const size_t i = base + (size_t)u * CP_THREADS;
        if (i < n) v[u] = X[i];
    }
#pragma unroll
    for (int u = 0; u < CP_UNROLL; ++u) {
        const size_t i = base + (size_t)u * CP_THREADS;
        if (i < n) Y[i] = v[u];
    }
}

// mode 0: x = X[i] (component filter); mode 1: x = (hi + lo) - B, also sums B^2
__global__ void __launch_bounds__(RED_THREADS)
k_reduce_pass1(const double *__restrict__ X, size_t n, int nvar, int comp, double *__restrict__ part)
{
    double s = 0.0, m = 0.0;
    for (size_t i = blockIdx.x * (size_t)RED_THREADS + threadIdx.x; i < n; i += (size_t)gridDim.x * RED_THREADS) {
        if (comp >= 0 && (int)(i % nvar) != comp) continue;
        double v = X[i];
        s += v * v;
        m = nmax(m, fabs(v));
    }
    block_reduce(s, m);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s; part[2 * blockIdx.x + 1] = m; }
}

__global__ void __launch_bounds__(RED_THREADS)
k_diff_pass1(const float *__restrict__ HI, const float *__restrict__ LO, const double *__restrict__ B, size_t n,
             double *__restrict__ part)
{
    double s = 0.0, m = 0.0, sb = 0.0, mb = 0.0;
    for (size_t i = blockIdx.x * (size_t)RED_THREADS + threadIdx.x; i < n; i += (size_t)gridDim.x * RED_THREADS) {
        double a = (double)HI[i] + (LO ? (double)LO[i] : 0.0);
        double d = a - B[i];
        s += d * d;
        m = nmax(m, fabs(d));
        sb += B[i] * B[i];
    }
    block_reduce(s, m);
    __syncthreads();
    block_reduce(sb, mb);
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = s; part[3 * blockIdx.x + 1] = m; part[3 * blockIdx.x + 2] = sb; }
}

__global__ void __launch_bounds__(RED_THREADS)
k_energy_pass1(const double *__restrict__ U, size_t ncell, double invK, double rho0, double *__restrict__ part)
{
    double s = 0.0, m = 0.0;
    for (size_t i = blockIdx.x * (size_t)RED_THREADS + threadIdx.x; i < ncell; i += (size_t)gridDim.x * RED_THREADS) {
        const double2 *p = reinterpret_cast<const double2 *>(U + 4 * i);
        double2 a = p[0], b = p[1];
        double e = a.x * a.x * invK + rho0 * (a.y * a.y + b.x * b.x + b.y * b.y);
        s += e;
        m = nmax(m, e);
    }
    block_reduce(s, m);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s; part[2 * blockIdx.x + 1] = m; }
}

__global__ void __launch_bounds__(RED_THREADS)
k_reduce_pass2(const double *__restrict__ part, int nblk, int stride, double *__restrict__ out)
{
    // fixed-order accumulation of the block partials (columns 0: sum, 1: max, 2: sum)
    double s = 0.0, m = 0.0, s2 = 0.0, dummy = 0.0;
    for (int i = threadIdx.x; i < nblk; i += RED_THREADS) {
        s += part[stride * i];
        m = nmax(m, part[stride * i + 1]);
        if (stride > 2) s2 += part[stride * i + 2];
    }
    block_reduce(s, m);
    __syncthreads();
    block_reduce(s2, dummy);
    if (threadIdx.x == 0) { out[0] = s; out[1] = m; out[2] = s2; }
}

// Partials (3 per block) + result of one reduction.
constexpr size_t RED_SCRATCH = 3 * RED_BLOCKS + 4;

// The partials of one call, taken from the stream-ordered pool of `s` and
// returned to it when the call ends, so calls on different streams never
// share them.
static int red_alloc(double **part, hipStream_t s)
{
    return hip_check(hipMallocAsync((void **)part, RED_SCRATCH * sizeof(double), s), "reduce scratch");
}

static int finish(double *part, int stride, double *host_out, int nout, hipStream_t s)
{
    double *res = part + 3 * RED_BLOCKS;
    hipLaunchKernelGGL(k_reduce_pass2, dim3(1), dim3(RED_THREADS), 0, s, part, RED_BLOCKS, stride, res);
    int rc = hip_check(hipGetLastError(), "reduce pass 2");
    double tmp[3];
    if (!rc) rc = hip_check(hipMemcpyAsync(tmp, res, 3 * sizeof(double), hipMemcpyDeviceToHost, s), "reduce copy");
    const int rf = hip_check(hipFreeAsync(part, s), "reduce scratch free");
    if (!rc) rc = rf;
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "reduce sync");
    if (rc) return rc;
    memcpy(host_out, tmp, nout * sizeof(double));
    return ACX_OK;
}

// ---- CTCS leapfrog (3d/ctcs/point_source/wave.c:138-215) -----------------

__device__ __forceinline__ double ctcs_source(double t, double f0)
{
    const double t0 = 4.0 / f0;   // wave.c:234-240
    return -2.0 * (t - t0) * f0 * f0 * exp(-1.0 * f0 * f0 * (t - t0) * (t - t0));
}

template <int ST>
__global__ void __launch_bounds__(256)
k_ctcs(const double *__restrict__ Pold, const double *__restrict__ P, double *__restrict__ Pnew,
       const double *__restrict__ C, acx_ctcs c, int it, double *__restrict__ rcv)
{
    const int nx = c.n[0], ny = c.n[1], nz = c.n[2];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nx) return;
    // periodic wrap of offsets within 2 cells (axes have 1 or at least 5
    // cells); an axis of 1 cell (2D / 1D programs) maps every offset to 0
    auto W = [](int a, int n) { return n == 1 ? 0 : (a < 0 ? a + n : (a >= n ? a - n : a)); };
    auto at = [&](int a, int b, int e) { return P[((size_t)W(e, nz) * ny + W(b, ny)) * nx + W(a, nx)]; };
    const double invdx2 = 1.0 / (c.dx * c.dx), invdy2 = 1.0 / (c.dy * c.dy), invdz2 = 1.0 / (c.dz * c.dz);
    const double p0 = at(i, j, k);
    double pxx, pyy, pzz;
    if (ST == 1) {
        pxx = (at(i + 1, j, k) - 2.0 * p0 + at(i - 1, j, k)) * invdx2;
        pyy = (at(i, j + 1, k) - 2.0 * p0 + at(i, j - 1, k)) * invdy2;
        pzz = (at(i, j, k + 1) - 2.0 * p0 + at(i, j, k - 1)) * invdz2;
    } else {
        pxx = (-1.0 / 12 * at(i + 2, j, k) + 4.0 / 3.0 * at(i + 1, j, k) - 5.0 / 2.0 * p0 + 4.0 / 3.0 * at(i - 1, j, k) -
               1.0 / 12.0 * at(i - 2, j, k)) * invdx2;
        pyy = (-1.0 / 12 * at(i, j + 2, k) + 4.0 / 3.0 * at(i, j + 1, k) - 5.0 / 2.0 * p0 + 4.0 / 3.0 * at(i, j - 1, k) -
               1.0 / 12.0 * at(i, j - 2, k)) * invdy2;
        pzz = (-1.0 / 12 * at(i, j, k + 2) + 4.0 / 3.0 * at(i, j, k + 1) - 5.0 / 2.0 * p0 + 4.0 / 3.0 * at(i, j, k - 1) -
               1.0 / 12.0 * at(i, j, k - 2)) * invdz2;
    }
    const size_t o = ((size_t)k * ny + j) * nx + i;
    const double cc = C[o], dt = c.dt;
    double pn = 2.0 * p0 - Pold[o] + cc * cc * dt * dt * (pxx + pyy + pzz);
    if (i == c.src[0] && j == c.src[1] && k == c.src[2]) pn += ctcs_source(it * dt, c.f0) / (c.dx * c.dy * c.dz) * dt * dt;
    if (rcv && i == c.rcv[0] && j == c.rcv[1] && k == c.rcv[2]) rcv[0] = p0;
    Pnew[o] = pn;
}

// The same update as a 2.5D z-march for 3D grids (every axis >= 5 cells): a
// block owns a 32 x 8 column tile and a chunk of zc planes; each thread
// keeps its column's planes z-ST .. z+ST in registers (the z part of the
// Laplacian) and plane z sits in LDS with an ST-cell x/y halo (double
// buffered, one barrier per plane), so P is read about once per update
// instead of through 13 scattered loads.  Arithmetic as in k_ctcs, term by
// term.  zc: planes per block (32, fewer for small grids, see the launcher).
constexpr int CT_BX = 32, CT_BY = 8;

template <int ST>
__global__ void __launch_bounds__(CT_BX *CT_BY)
k_ctcs_z(const double *__restrict__ Pold, const double *__restrict__ P, double *__restrict__ Pnew,
         const double *__restrict__ C, acx_ctcs c, int it, int zc, double *__restrict__ rcv)
{
    constexpr int LX = CT_BX + 2 * ST, LY = CT_BY + 2 * ST, NH = 2 * ST * (CT_BX + CT_BY);
    static_assert(NH <= CT_BX * CT_BY, "every halo cell of a plane needs a thread of its own");
    __shared__ double pl[2][LY][LX];
    const int nx = c.n[0], ny = c.n[1], nz = c.n[2];
    const int ntx = (nx + CT_BX - 1) / CT_BX, nty = (ny + CT_BY - 1) / CT_BY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * CT_BX, y0 = ((b / ntx) % nty) * CT_BY, zb = (b / (ntx * nty)) * zc;
    const int ze = min(zb + zc, nz);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * CT_BX + tx;
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i < nx && j < ny;
    auto M = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };   // set-up only
    auto Wz = [nz](int a) { return a < 0 ? a + nz : (a >= nz ? a - nz : a); };   // |offset| <= ST + 1 < nz
    const size_t plane = (size_t)nx * ny;
    // columns past the grid edge hold the periodic image, so the last tile's
    // LDS row also serves as the wrapped x/y neighbours of the edge cells
    const size_t col = (size_t)M(j, ny) * nx + M(i, nx);
    int hx = -1, hy = -1;
    size_t hcol = 0;
    if (tid < NH) {
        int t = tid;
        if (t < 2 * ST * CT_BY) {
            const int row = t / (2 * ST), k = t % (2 * ST);
            hy = row + ST;
            hx = k < ST ? k : CT_BX + k;
        } else {
            t -= 2 * ST * CT_BY;
            const int cx = t % CT_BX, k = t / CT_BX;
            hx = cx + ST;
            hy = k < ST ? k : CT_BY + k;
        }
        hcol = (size_t)M(y0 + hy - ST, ny) * nx + M(x0 + hx - ST, nx);
    }
    const double invdx2 = 1.0 / (c.dx * c.dx), invdy2 = 1.0 / (c.dy * c.dy), invdz2 = 1.0 / (c.dz * c.dz);
    const double dt = c.dt;
    double q[2 * ST + 1];
#pragma unroll
    for (int d = -ST; d <= ST; ++d) q[d + ST] = P[(size_t)Wz(zb + d) * plane + col];
    int buf = 0;
    for (int z = zb; z < ze; ++z) {
        const double nxt = P[(size_t)Wz(z + ST + 1) * plane + col];   // enters the queue at the end
        const size_t o = (size_t)z * plane + col;
        const double po = Pold[o], cc = C[o];
        pl[buf][ty + ST][tx + ST] = q[ST];
        if (hx >= 0) pl[buf][hy][hx] = P[(size_t)z * plane + hcol];
        __syncthreads();
        if (active) {
            const double p0 = q[ST];
            const double *row = &pl[buf][ty + ST][tx + ST];
            double pxx, pyy, pzz;
            if (ST == 1) {
                pxx = (row[1] - 2.0 * p0 + row[-1]) * invdx2;
                pyy = (pl[buf][ty + ST + 1][tx + ST] - 2.0 * p0 + pl[buf][ty + ST - 1][tx + ST]) * invdy2;
                pzz = (q[2] - 2.0 * p0 + q[0]) * invdz2;
            } else {
                pxx = (-1.0 / 12 * row[2] + 4.0 / 3.0 * row[1] - 5.0 / 2.0 * p0 + 4.0 / 3.0 * row[-1] -
                       1.0 / 12.0 * row[-2]) * invdx2;
                pyy = (-1.0 / 12 * pl[buf][ty + ST + 2][tx + ST] + 4.0 / 3.0 * pl[buf][ty + ST + 1][tx + ST] -
                       5.0 / 2.0 * p0 + 4.0 / 3.0 * pl[buf][ty + ST - 1][tx + ST] -
                       1.0 / 12.0 * pl[buf][ty + ST - 2][tx + ST]) * invdy2;
                pzz = (-1.0 / 12 * q[2 * ST] + 4.0 / 3.0 * q[ST + 1] - 5.0 / 2.0 * p0 + 4.0 / 3.0 * q[ST - 1] -
                       1.0 / 12.0 * q[0]) * invdz2;
            }
            double pn = 2.0 * p0 - po + cc * cc * dt * dt * (pxx + pyy + pzz);
            if (i == c.src[0] && j == c.src[1] && z == c.src[2])
                pn += ctcs_source(it * dt, c.f0) / (c.dx * c.dy * c.dz) * dt * dt;
            if (rcv && i == c.rcv[0] && j == c.rcv[1] && z == c.rcv[2]) rcv[0] = p0;
            Pnew[o] = pn;
        }
#pragma unroll
        for (int d = 0; d < 2 * ST; ++d) q[d] = q[d + 1];
        q[2 * ST] = nxt;
        buf ^= 1;
    }
}

// The same update on an LDS-DMA march (3D grids with an even x extent and
// 16-byte aligned arrays): a block of 32 x 16 threads owns a 32 x 16 column
// tile; each plane of P with a 2-cell x/y halo (36 x 20 cells) and the
// tile's Pold and C go straight from memory into one slot of a 5-slot LDS
// ring with global_load_lds_dwordx4 (16 bytes = two x-neighbours per lane).
// At iteration z the slots hold planes z-1 .. z+3: plane z feeds the x/y
// stencils and the update, z+1 and z+2 have landed (z+2 enters the column's
// z-queue of planes z-2 .. z+2 in registers), z+3 and the plane z+4 issued in
// this iteration are in flight.  Loads complete in issue order, so
// s_waitcnt vmcnt(nj) (nj = this wavefront's loads of one plane) means plane
// z+2 has landed in this wavefront's part; one s_barrier then publishes all
// parts and frees the slot of plane z-1 for plane z+4.  The loads are issued
// from inline assembly so that the compiler does not guard every LDS read
// with a wait for all of them.  69.8 KB of LDS per block: two blocks per
// CU.  The arithmetic is k_ctcs_z's term by term.
struct CtDma {
    static constexpr int BX = 32, BY = 16, R = 2;                 // tile, halo loaded (covers ST 1 and 2)
    static constexpr int RW = BX + 2 * R, RH = BY + 2 * R;        // 36 x 20 cells of P
    static constexpr int NP = RW / 2 * RH;                        // 360 entries (cell pairs) of P
    static constexpr int NI = BX / 2 * BY;                        // 256 entries of Pold, and of C
    static constexpr int NENT = NP + 2 * NI;                      // 872
    static constexpr int SLOT = NENT * 16;                        // 13952 bytes
    static constexpr int NSLOT = 5;
    static constexpr int NWAVE = BX * BY / 64;                    // 8
    static constexpr int NCHUNK = (NENT + 63) / 64;               // 14 wavefront loads per plane
    static constexpr int MAXJ = (NCHUNK + NWAVE - 1) / NWAVE;     // 2
};

template <int ST>
__global__ void __launch_bounds__(CtDma::BX *CtDma::BY) __attribute__((amdgpu_waves_per_eu(4)))
k_ctcs_dma(const double *__restrict__ Pold, const double *__restrict__ P, double *__restrict__ Pnew,
           const double *__restrict__ C, acx_ctcs c, int it, int zc, double *__restrict__ rcv)
{
    using D = CtDma;
    __shared__ __attribute__((aligned(16))) unsigned char ring[D::NSLOT * D::SLOT];
    const int nx = c.n[0], ny = c.n[1], nz = c.n[2];
    const int ntx = (nx + D::BX - 1) / D::BX, nty = (ny + D::BY - 1) / D::BY;
    const int b = xcd_logical_tile(blockIdx.x, gridDim.x);
    const int x0 = (b % ntx) * D::BX, y0 = ((b / ntx) % nty) * D::BY, zb = (b / (ntx * nty)) * zc;
    const int ze = min(zb + zc, nz);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * D::BX + tx;
    const int wave = tid >> 6, lane = tid & 63;
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i < nx && j < ny;
    const size_t plane = (size_t)nx * ny;
    auto M = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };

    // this lane's entries: chunk q = wave + 8 k covers entries 64 q .. 64 q + 63
    const int nj = wave + D::NWAVE * (D::MAXJ - 1) < D::NCHUNK ? D::MAXJ : D::MAXJ - 1;
    const double *eptr[D::MAXJ];   // P, Pold or C at the entry's cell of plane 0
    bool eok[D::MAXJ];
#pragma unroll
    for (int k = 0; k < D::MAXJ; ++k) {
        const int e = (wave + D::NWAVE * k) * 64 + lane;
        eok[k] = k < nj && e < D::NENT;
        int gx, gy;
        const double *base;
        if (e < D::NP) {
            const int r = e / (D::RW / 2), cp = e - r * (D::RW / 2);
            gx = M(x0 - D::R + 2 * cp, nx);
            gy = M(y0 - D::R + r, ny);
            base = P;
        } else {
            const int e2 = (e - D::NP) % D::NI, r = e2 / (D::BX / 2), cp = e2 - r * (D::BX / 2);
            gx = M(x0 + 2 * cp, nx);
            gy = M(y0 + r, ny);
            base = e < D::NP + D::NI ? Pold : C;
        }
        eptr[k] = base + (eok[k] ? (size_t)gy * nx + gx : 0);
    }
    auto slot_of = [&](int z) { return (z - zb + 2) % D::NSLOT; };
    auto issue = [&](int z, int slot) {
        const size_t zo = (size_t)M(z, nz) * plane;
        unsigned char *dst = ring + slot * D::SLOT;
#pragma unroll
        for (int k = 0; k < D::MAXJ; ++k) {
            if (k < nj) {
                const unsigned lds = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(
                    (__attribute__((address_space(3))) unsigned char *)(dst + (wave + D::NWAVE * k) * 1024)));
                const double *g = eptr[k] + zo;
                if (eok[k])
                    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                                 :: "s"(lds), "v"(g) : "memory", "m0");
            }
        }
    };
    // vmcnt (gfx9 encoding, expcnt 7, lgkmcnt 15): all loads, or all but the
    // nj of the most recent plane
    auto wait_planes = [&](bool keep_last) {
        if (!keep_last) __builtin_amdgcn_s_waitcnt(0x0F70);
        else if (nj == 2) __builtin_amdgcn_s_waitcnt(0x0F72);
        else __builtin_amdgcn_s_waitcnt(0x0F71);
    };
    // P at region row r, column cc of a slot; Pold / C of the tile cell (ty, tx)
    auto pv = [&](int slot, int r, int cc) {
        return *reinterpret_cast<const double *>(ring + slot * D::SLOT + (r * D::RW + cc) * 8);
    };
    auto tile_v = [&](int slot, int part) {
        return *reinterpret_cast<const double *>(ring + slot * D::SLOT + D::NP * 16 + part * D::NI * 16 +
                                                 (ty * D::BX + tx) * 8);
    };

    // prologue: planes zb-2 .. zb+2 into slots 0 .. 4, the column queue, then zb+3
    for (int d = 0; d < D::NSLOT; ++d) issue(zb - 2 + d, d);
    wait_planes(false);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    double q[5];
#pragma unroll
    for (int d = 0; d < 5; ++d) q[d] = pv(d, ty + D::R, tx + D::R);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (zb + 3 <= ze + 1) issue(zb + 3, slot_of(zb + 3));

    const double invdx2 = 1.0 / (c.dx * c.dx), invdy2 = 1.0 / (c.dy * c.dy), invdz2 = 1.0 / (c.dz * c.dz);
    const double dt = c.dt;
    const size_t col = (size_t)M(j, ny) * nx + M(i, nx);
    for (int z = zb; z < ze; ++z) {
        if (z > zb) {
            // plane z+2 landed: only plane z+3's loads may be outstanding
            wait_planes(z + 3 <= ze + 1);
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
#pragma unroll
            for (int d = 0; d < 4; ++d) q[d] = q[d + 1];
            q[4] = pv(slot_of(z + 2), ty + D::R, tx + D::R);
        }
        // every wavefront is past iteration z-1: its slot takes plane z+4
        if (z + 4 <= ze + 1) issue(z + 4, slot_of(z + 4));
        const int s0 = slot_of(z);
        if (active) {
            const double p0 = q[2];
            const double po = tile_v(s0, 0), cc = tile_v(s0, 1);
            const int r0 = ty + D::R, c0 = tx + D::R;
            double pxx, pyy, pzz;
            if (ST == 1) {
                pxx = (pv(s0, r0, c0 + 1) - 2.0 * p0 + pv(s0, r0, c0 - 1)) * invdx2;
                pyy = (pv(s0, r0 + 1, c0) - 2.0 * p0 + pv(s0, r0 - 1, c0)) * invdy2;
                pzz = (q[3] - 2.0 * p0 + q[1]) * invdz2;
            } else {
                pxx = (-1.0 / 12 * pv(s0, r0, c0 + 2) + 4.0 / 3.0 * pv(s0, r0, c0 + 1) - 5.0 / 2.0 * p0 +
                       4.0 / 3.0 * pv(s0, r0, c0 - 1) - 1.0 / 12.0 * pv(s0, r0, c0 - 2)) * invdx2;
                pyy = (-1.0 / 12 * pv(s0, r0 + 2, c0) + 4.0 / 3.0 * pv(s0, r0 + 1, c0) - 5.0 / 2.0 * p0 +
                       4.0 / 3.0 * pv(s0, r0 - 1, c0) - 1.0 / 12.0 * pv(s0, r0 - 2, c0)) * invdy2;
                pzz = (-1.0 / 12 * q[4] + 4.0 / 3.0 * q[3] - 5.0 / 2.0 * p0 + 4.0 / 3.0 * q[1] -
                       1.0 / 12.0 * q[0]) * invdz2;
            }
            double pn = 2.0 * p0 - po + cc * cc * dt * dt * (pxx + pyy + pzz);
            if (i == c.src[0] && j == c.src[1] && z == c.src[2])
                pn += ctcs_source(it * dt, c.f0) / (c.dx * c.dy * c.dz) * dt * dt;
            if (rcv && i == c.rcv[0] && j == c.rcv[1] && z == c.rcv[2]) rcv[0] = p0;
            Pnew[(size_t)z * plane + col] = pn;
        }
    }
    // every load issued has been waited for by the last iteration (vmcnt 0)
}

// ---- host fields ---------------------------------------------------------

static inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline double unif(uint64_t &s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// 1D cell averages of exp(-(x-c)^2/(2 sigma^2)) on the periodic unit interval
static void gauss_1d(int n, double h, double c, double sigma, std::vector<double> &out)
{
    out.assign(n, 0.0);
    const double k = sigma * sqrt(M_PI / 2.0) / h, r = 1.0 / (sigma * sqrt(2.0));
    for (int i = 0; i < n; ++i) {
        const double a = i * h, b = (i + 1) * h;
        double acc = 0.0;
        for (int img = -1; img <= 1; ++img) acc += erf((b - c + img) * r) - erf((a - c + img) * r);
        out[i] = k * acc;
    }
}

} // namespace acx

using namespace acx;

extern "C" {

int acx_norms_f64(const double *X, size_t ncells, int nvar, int comp, double *out, void *stream)
{
    if (!X || !out || nvar < 1 || comp >= nvar) return set_error(ACX_EINVAL, "norms: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    int rc = red_alloc(&part, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_pass1, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, X, ncells * nvar, nvar, comp, part);
    if ((rc = hip_check(hipGetLastError(), "norms pass 1"))) {
        (void)hipFreeAsync(part, s);
        return rc;
    }
    double r[3];
    if ((rc = finish(part, 2, r, 3, s))) return rc;
    out[0] = r[0];
    out[1] = r[1];
    return ACX_OK;
}

int acx_vec_copy_f64(const double *X, double *Y, size_t n, void *stream)
{
    if (n == 0) return ACX_OK;
    if (!X || !Y) return set_error(ACX_EINVAL, "vec copy: null pointer");
    if ((X < Y + n) && (Y < X + n) && X != Y) return set_error(ACX_EINVAL, "vec copy: overlapping arrays");
    if (X == Y) return ACX_OK;
    hipStream_t s = (hipStream_t)stream;
    const size_t per_block = (size_t)CP_THREADS * CP_UNROLL;
    const bool aligned = (((uintptr_t)X | (uintptr_t)Y) & 15) == 0;
    const size_t nv = aligned ? n / 2 : n;                 // elements of the main launch
    const size_t blocks = (nv + per_block - 1) / per_block;
    if (blocks > 0x7fffffffULL) return set_error(ACX_EINVAL, "vec copy: %zu elements is too many", n);
    if (blocks) {
        if (aligned)
            hipLaunchKernelGGL(k_copy2, dim3((unsigned)blocks), dim3(CP_THREADS), 0, s, (const acx_dbl2 *)X,
                               (acx_dbl2 *)Y, nv);
        else
            hipLaunchKernelGGL(k_copy1, dim3((unsigned)blocks), dim3(CP_THREADS), 0, s, X, Y, nv);
    }
    if (aligned && (n & 1)) hipLaunchKernelGGL(k_copy1, dim3(1), dim3(CP_THREADS), 0, s, X + n - 1, Y + n - 1, (size_t)1);
    return hip_check(hipGetLastError(), "vec copy");
}

int acx_diff_f32c_f64(const float *HI, const float *LO, const double *B, size_t n, double *out, void *stream)
{
    if (!HI || !B || !out) return set_error(ACX_EINVAL, "diff: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    int rc = red_alloc(&part, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_diff_pass1, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, HI, LO, B, n, part);
    if ((rc = hip_check(hipGetLastError(), "diff pass 1"))) {
        (void)hipFreeAsync(part, s);
        return rc;
    }
    double r[3];
    if ((rc = finish(part, 3, r, 3, s))) return rc;
    out[0] = r[2] > 0 ? sqrt(r[0] / r[2]) : sqrt(r[0]);
    out[1] = r[1];
    return ACX_OK;
}

int acx_energy_f64(const acx_grid *g, const double *U, double *out, void *stream)
{
    if (!g || !U || !out || g->system != ACX_ACOUSTIC4) return set_error(ACX_EINVAL, "energy: ACOUSTIC4 state required");
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    int rc = red_alloc(&part, s);
    if (rc) return rc;
    const size_t nc = (size_t)g->n[0] * g->n[1] * g->n[2];
    const double K = g->rho0 * g->c0 * g->c0;
    hipLaunchKernelGGL(k_energy_pass1, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, U, nc, 1.0 / K, g->rho0, part);
    if ((rc = hip_check(hipGetLastError(), "energy pass 1"))) {
        (void)hipFreeAsync(part, s);
        return rc;
    }
    double r[3];
    if ((rc = finish(part, 2, r, 3, s))) return rc;
    out[0] = 0.5 * r[0] * g->h * g->h * (g->n[2] == 1 ? 1.0 : g->h);   // cell area on a 2D grid
    return ACX_OK;
}

int acx_ctcs_update_f64(const acx_ctcs *c, const double *POLD, const double *P, double *PNEW, const double *C,
                        int it, double *rcv_out, void *stream)
{
    if (!c || !POLD || !P || !PNEW || !C || PNEW == P) return set_error(ACX_EINVAL, "ctcs: bad arguments");
    for (int d = 0; d < 3; ++d)
        if (c->n[d] < 5 && c->n[d] != 1) return set_error(ACX_EINVAL, "ctcs: at least 5 cells per axis (or 1)");
    if (c->stencil != 1 && c->stencil != 2) return set_error(ACX_EINVAL, "stencil width can be either 1 or 2");
    hipStream_t s = (hipStream_t)stream;
    // the receiver sample goes through a one-value buffer of this call (none
    // when the caller does not ask for it)
    double *rcv = nullptr;
    if (rcv_out) {
        const int ra = hip_check(hipMallocAsync((void **)&rcv, sizeof(double), s), "ctcs receiver buffer");
        if (ra) return ra;
    }
    const bool aligned = ((((uintptr_t)POLD) | ((uintptr_t)P) | ((uintptr_t)C)) & 15) == 0;
    if (c->n[0] >= 5 && c->n[1] >= 5 && c->n[2] >= 5 && c->n[0] % 2 == 0 && aligned &&
        (size_t)c->n[0] * c->n[1] < 0xffffffffu) {
        // 3D, LDS-DMA march: z-chunk from the round model, two blocks per CU
        // (measured at 256^3: 144 -> 152 G cell-updates/s, bit-identical to
        // the register-queue march below)
        const int ntx = (c->n[0] + CtDma::BX - 1) / CtDma::BX, nty = (c->n[1] + CtDma::BY - 1) / CtDma::BY;
        const int zc = auto_zc(ntx, nty, c->n[2], device_cus() * 2, 4);
        const size_t nb = (size_t)ntx * nty * ((c->n[2] + zc - 1) / zc);
        dim3 blk(CtDma::BX, CtDma::BY), grd((unsigned)nb);
        if (c->stencil == 1) hipLaunchKernelGGL(k_ctcs_dma<1>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, zc, rcv);
        else hipLaunchKernelGGL(k_ctcs_dma<2>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, zc, rcv);
    } else if (c->n[0] >= 5 && c->n[1] >= 5 && c->n[2] >= 5) {
        // 3D: z-march; 32 planes per block, halved (down to 4) below 1024 blocks
        const size_t cols = (size_t)((c->n[0] + CT_BX - 1) / CT_BX) * ((c->n[1] + CT_BY - 1) / CT_BY);
        int zc = 32;
        while (zc > 4 && cols * ((c->n[2] + zc - 1) / zc) < 1024) zc /= 2;
        const size_t nb = cols * ((c->n[2] + zc - 1) / zc);
        dim3 blk(CT_BX, CT_BY), grd((unsigned)nb);
        if (c->stencil == 1) hipLaunchKernelGGL(k_ctcs_z<1>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, zc, rcv);
        else hipLaunchKernelGGL(k_ctcs_z<2>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, zc, rcv);
    } else {
        // 2D / 1D programs (axes of one cell): one thread per cell
        dim3 blk(256), grd((c->n[0] + 255) / 256, c->n[1], c->n[2]);
        if (c->stencil == 1) hipLaunchKernelGGL(k_ctcs<1>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, rcv);
        else hipLaunchKernelGGL(k_ctcs<2>, grd, blk, 0, s, POLD, P, PNEW, C, *c, it, rcv);
    }
    int rc = hip_check(hipGetLastError(), "ctcs launch");
    if (rcv_out) {
        if (!rc) rc = hip_check(hipMemcpyAsync(rcv_out, rcv, sizeof(double), hipMemcpyDeviceToHost, s), "ctcs rcv copy");
        const int rf = hip_check(hipFreeAsync(rcv, s), "ctcs receiver buffer free");
        if (!rc) rc = rf;
        if (!rc) rc = hip_check(hipStreamSynchronize(s), "ctcs sync");
    }
    return rc;
}

int acx_field_gaussians(const acx_grid *g, uint64_t seed, int npulse, double amin, double amax, double smin,
                        double smax, double *U)
{
    if (!g || !U || npulse < 0) return set_error(ACX_EINVAL, "field: bad arguments");
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2], nv = g->system == ACX_ACOUSTIC4 ? 4 : 2;
    const size_t nc = (size_t)nx * ny * nz;
    memset(U, 0, nc * nv * sizeof(double));
    uint64_t st = seed;
    std::vector<double> fx, fy, fz;
    for (int q = 0; q < npulse; ++q) {
        const double A = amin + (amax - amin) * unif(st);
        const double sg = smin + (smax - smin) * unif(st);
        const double cx = unif(st), cy = unif(st), cz = unif(st);
        gauss_1d(nx, 1.0 / nx, cx, sg, fx);
        gauss_1d(ny, 1.0 / ny, cy, sg, fy);
        gauss_1d(nz, 1.0 / nz, cz, sg, fz);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                const double a = A * fz[k] * fy[j];
                double *row = U + ((size_t)k * ny + j) * nx * nv;
                for (int i = 0; i < nx; ++i) row[(size_t)i * nv] += a * fx[i];
            }
    }
    return ACX_OK;
}

int acx_field_plane_wave(const acx_grid *g, const int k[3], double A, double t, double *U)
{
    if (!g || !U || g->system != ACX_ACOUSTIC4) return set_error(ACX_EINVAL, "plane wave: ACOUSTIC4 grid required");
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const double kk = sqrt((double)k[0] * k[0] + (double)k[1] * k[1] + (double)k[2] * k[2]);
    if (kk == 0) return set_error(ACX_EINVAL, "plane wave: k = 0");
    const double w = 2.0 * M_PI * kk * g->c0;
    // cell averages: point value at the centre times prod sinc(pi k_d h_d)
    auto sinc = [](double x) { return x == 0 ? 1.0 : sin(x) / x; };
    const double damp = sinc(M_PI * k[0] / nx) * sinc(M_PI * k[1] / ny) * sinc(M_PI * k[2] / nz);
    const double ap = A * cos(w * t) * damp, au = -(A / (g->rho0 * g->c0)) * sin(w * t) * damp;
    for (int c = 0; c < nz; ++c)
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                const double th = 2.0 * M_PI * (k[0] * (a + 0.5) / nx + k[1] * (b + 0.5) / ny + k[2] * (c + 0.5) / nz);
                double *o = U + (((size_t)c * ny + b) * nx + a) * 4;
                const double sn = sin(th), cs = cos(th);
                o[0] = ap * sn;
                o[1] = au * (k[0] / kk) * cs;
                o[2] = au * (k[1] / kk) * cs;
                o[3] = au * (k[2] / kk) * cs;
            }
    return ACX_OK;
}

} // extern "C"
