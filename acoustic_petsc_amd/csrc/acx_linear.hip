// acx_linear.hip -- fused RHS + SSP-stage kernel for the linear (shipped)
// CWENO4 scheme, gamma_hi = 1, gamma_lo = 0.
//
// With those weights the reference's weno() returns the 4th-order central
// polynomial exactly (3d/weno4/plane_wave/reconstruction.c:42-59,236-257:
// gamma_R3 = 0 so w_R3 = 0 and wt_ratio = 1), and the 2x2 Gauss-point face
// averages of its value / normal gradient depend only on 1D five-point
// combinations along the face normal.  The RHS is then an axis-aligned
// 7-point star stencil per axis (see acx_device.h), fused here with the
// SSP stage combination  OUT = a*U0 + b*Q + c*F(Q).
//
// Layout: 2.5D blocking.  A block owns a BX x BY column tile and marches a
// chunk of z-planes.  Each thread keeps its column's z-neighbourhood
// (z-3..z+3) in registers and prefetches z+4; the current plane with an
// x/y halo of 3 is staged in LDS (double-buffered, one barrier per plane).
// HBM traffic per stage is one read of Q (+ U0) and one write of OUT; halo
// and z-chunk overlap re-reads are served by L2 / Infinity Cache.
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

template <typename T, int SYS, int BX, int BY>
struct LinTile {
    static constexpr int NV = SYS == ACX_ACOUSTIC4 ? 4 : 2;   // components in memory
    static constexpr int NS = SYS == ACX_ACOUSTIC4 ? 4 : 1;   // components in the stencil
    static constexpr int LX = BX + 6, LY = BY + 6;
    static constexpr int NHALO = 6 * BX + 6 * BY;              // star stencil: no plane corners
};

// Epilogue kinds
enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3 };

template <typename T, int SYS, int BX, int BY, int EPI>
__global__ void __launch_bounds__(BX *BY)
k_linear(LinArgs p)
{
    using Tile = LinTile<T, SYS, BX, BY>;
    constexpr int NV = Tile::NV, NS = Tile::NS, LX = Tile::LX, LY = Tile::LY;
    __shared__ T lds[2][NS][LY][LX];

    const T *__restrict__ Q = static_cast<const T *>(p.q);
    const T *__restrict__ D = static_cast<const T *>(p.d);       // F32C: increment added to Q
    const T *__restrict__ U0 = static_cast<const T *>(p.u0);
    const T *__restrict__ LO = static_cast<const T *>(p.lo);
    T *__restrict__ OUT = static_cast<T *>(p.out);
    T *__restrict__ OUT2 = static_cast<T *>(p.out2);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int x0 = blockIdx.x * BX, y0 = blockIdx.y * BY;
    const int zb = blockIdx.z * p.zc;
    const int ze = min(zb + p.zc, nz);
    const bool active = (x0 + tx < nx) && (y0 + ty < ny);
    const size_t plane = (size_t)nx * ny;

    // own column
    float sgx = 1.f, sgy = 1.f;
    const int mx = map_axis(x0 + tx, nx, p.bc[0] == ACX_PERIODIC, sgx);
    const int my = map_axis(y0 + ty, ny, p.bc[1] == ACX_PERIODIC, sgy);
    const size_t col = (size_t)my * nx + mx;
    const float csign = sgx * sgy;

    // halo cell owned by this thread (x halo for rows 0..BY-1, y halo for cols 0..BX-1)
    int hx = -1, hy = -1;
    size_t hcol = 0;
    float hsign = 1.f;
    if (tid < Tile::NHALO) {
        int t = tid;
        if (t < 6 * BY) {            // x halo: 3 left, 3 right per row
            int row = t / 6, k = t % 6;
            hy = row + 3;
            hx = k < 3 ? k : BX + k;  // 0,1,2 | BX+3..BX+5
        } else {
            t -= 6 * BY;
            int colx = t % BX, k = t / BX;
            hx = colx + 3;
            hy = k < 3 ? k : BY + k;
        }
        float s1 = 1.f, s2 = 1.f;
        int gx = map_axis(x0 + hx - 3, nx, p.bc[0] == ACX_PERIODIC, s1);
        int gy = map_axis(y0 + hy - 3, ny, p.bc[1] == ACX_PERIODIC, s2);
        hcol = (size_t)gy * nx + gx;
        hsign = s1 * s2;
    }

    auto zmap = [&](int z, float &s) -> size_t {
        s = 1.f;
        return (size_t)map_axis(z, nz, p.bc[2] == ACX_PERIODIC, s) * plane;
    };
    // value of a cell as the stencil sees it (component 0 carries the ghost sign)
    auto fetch = [&](size_t idx, float sign) -> Cell<T, NV> {
        Cell<T, NV> c = load_cell<T, NV>(Q, idx);
        if constexpr (EPI == EPI_F32C) {
            if (D) {
                Cell<T, NV> d = load_cell<T, NV>(D, idx);
#pragma unroll
                for (int v = 0; v < NV; ++v) c.v[v] += d.v[v];
            }
        }
        c.v[0] *= T(sign);
        return c;
    };

    // z queue: planes z-3..z+3 of the own column, plus the prefetched z+4
    Cell<T, NV> qz[7];
#pragma unroll
    for (int d = -3; d <= 3; ++d) {
        float s;
        size_t off = zmap(zb + d, s);
        qz[d + 3] = fetch(off + col, s * csign);
    }
    Cell<T, NV> hv;
    if (hx >= 0) {
        float s;
        size_t off = zmap(zb, s);
        hv = fetch(off + hcol, s * hsign);
    }

    const T inv48h = T(p.inv48h), Kc = T(p.K), irho = T(p.invrho), sLLF = T(p.s);
    const T lap = T(p.lapc);   // WAVE2: c^2 / (48 h^2)

    int buf = 0;
    for (int z = zb; z < ze; ++z) {
        // prefetch the next plane of the column and of the halo cell
        Cell<T, NV> qn, hn;
        {
            float s;
            size_t off = zmap(z + 4, s);
            qn = fetch(off + col, s * csign);
            if (hx >= 0) {
                size_t off1 = zmap(z + 1, s);
                hn = fetch(off1 + hcol, s * hsign);
            }
        }
        // stage the current plane
#pragma unroll
        for (int v = 0; v < NS; ++v) lds[buf][v][ty + 3][tx + 3] = qz[3].v[v];
        if (hx >= 0) {
#pragma unroll
            for (int v = 0; v < NS; ++v) lds[buf][v][hy][hx] = hv.v[v];
        }
        __syncthreads();

        Cell<T, NV> r;
        if constexpr (SYS == ACX_ACOUSTIC4) {
            T accB[4], dAp[3], dAvel = T(0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T *row = &lds[buf][v][ty + 3][tx + 3];
                const T xm3 = row[-3], xm2 = row[-2], xm1 = row[-1], xp1 = row[1], xp2 = row[2], xp3 = row[3];
                const T ym3 = lds[buf][v][ty][tx + 3], ym2 = lds[buf][v][ty + 1][tx + 3], ym1 = lds[buf][v][ty + 2][tx + 3];
                const T yp1 = lds[buf][v][ty + 4][tx + 3], yp2 = lds[buf][v][ty + 5][tx + 3], yp3 = lds[buf][v][ty + 6][tx + 3];
                const T q0 = qz[3].v[v];
                accB[v] = stencil_B7(xm3, xm2, xm1, q0, xp1, xp2, xp3) +
                          stencil_B7(ym3, ym2, ym1, q0, yp1, yp2, yp3) +
                          stencil_B7(qz[0].v[v], qz[1].v[v], qz[2].v[v], q0, qz[4].v[v], qz[5].v[v], qz[6].v[v]);
                if (v == 0) {
                    dAp[0] = stencil_A7(xm3, xm2, xm1, xp1, xp2, xp3);
                    dAp[1] = stencil_A7(ym3, ym2, ym1, yp1, yp2, yp3);
                    dAp[2] = stencil_A7(qz[0].v[0], qz[1].v[0], qz[2].v[0], qz[4].v[0], qz[5].v[0], qz[6].v[0]);
                } else if (v == 1) {
                    dAvel += stencil_A7(xm3, xm2, xm1, xp1, xp2, xp3);
                } else if (v == 2) {
                    dAvel += stencil_A7(ym3, ym2, ym1, yp1, yp2, yp3);
                } else {
                    dAvel += stencil_A7(qz[0].v[3], qz[1].v[3], qz[2].v[3], qz[4].v[3], qz[5].v[3], qz[6].v[3]);
                }
            }
            r.v[0] = inv48h * (sLLF * accB[0] - Kc * dAvel);
            r.v[1] = inv48h * (sLLF * accB[1] - irho * dAp[0]);
            r.v[2] = inv48h * (sLLF * accB[2] - irho * dAp[1]);
            r.v[3] = inv48h * (sLLF * accB[3] - irho * dAp[2]);
        } else {
            const T *row = &lds[buf][0][ty + 3][tx + 3];
            const T q0 = qz[3].v[0];
            T L = stencil_L7(row[-3], row[-2], row[-1], q0, row[1], row[2], row[3]) +
                  stencil_L7(lds[buf][0][ty][tx + 3], lds[buf][0][ty + 1][tx + 3], lds[buf][0][ty + 2][tx + 3], q0,
                             lds[buf][0][ty + 4][tx + 3], lds[buf][0][ty + 5][tx + 3], lds[buf][0][ty + 6][tx + 3]) +
                  stencil_L7(qz[0].v[0], qz[1].v[0], qz[2].v[0], q0, qz[4].v[0], qz[5].v[0], qz[6].v[0]);
            r.v[0] = qz[3].v[1];     // du/dt = v (rhs_function.c:330)
            r.v[1] = lap * L;
        }

        if (active) {
            const size_t idx = (size_t)z * plane + col;   // owned cell: no ghost mapping
            Cell<T, NV> o;
            if constexpr (EPI == EPI_RHS) {
                o = r;
            } else if constexpr (EPI == EPI_AXPY) {
                const T cc = T(p.c);
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = qz[3].v[v] + cc * r.v[v];
            } else if constexpr (EPI == EPI_COMB) {
                const T aa = T(p.a), bb = T(p.b), cc = T(p.c);
                Cell<T, NV> u0 = load_cell<T, NV>(U0, idx);
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = aa * u0.v[v] + bb * qz[3].v[v] + cc * r.v[v];
            } else {
                // compensated FP32: increments relative to the step's base (HI + LO)
                const T bb = T(p.b), cc = T(p.c);
                Cell<T, NV> dcur;
                if (D) dcur = load_cell<T, NV>(D, idx);
                else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) dcur.v[v] = T(0);
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = bb * dcur.v[v] + cc * r.v[v];
                if (p.final_stage) {
                    Cell<T, NV> hi = load_cell<T, NV>(Q, idx), lo = load_cell<T, NV>(LO, idx), nh, nl;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        float t = (float)lo.v[v] + (float)o.v[v];
                        float s, e;
                        two_sum((float)hi.v[v], t, s, e);
                        nh.v[v] = s;
                        nl.v[v] = e;
                    }
                    store_cell<T, NV>(OUT2, idx, nl);
                    o = nh;
                }
            }
            store_cell<T, NV>(OUT, idx, o);
        }

        // shift the queue
#pragma unroll
        for (int k = 0; k < 6; ++k) qz[k] = qz[k + 1];
        qz[6] = qn;
        hv = hn;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------

template <typename T, int SYS, int EPI>
static int launch_linear_t(const LinArgs &a, hipStream_t s)
{
    constexpr int BX = 64, BY = 8;
    dim3 blk(BX, BY, 1);
    dim3 grd((a.nx + BX - 1) / BX, (a.ny + BY - 1) / BY, (a.nz + a.zc - 1) / a.zc);
    hipLaunchKernelGGL((k_linear<T, SYS, BX, BY, EPI>), grd, blk, 0, s, a);
    return hip_check(hipGetLastError(), "k_linear launch");
}

int launch_linear(const LinArgs &a, int sys, int epi, bool fp32, hipStream_t s)
{
    if (fp32) {
        if (sys == ACX_ACOUSTIC4) {
            if (epi == EPI_F32C) return launch_linear_t<float, ACX_ACOUSTIC4, EPI_F32C>(a, s);
        } else {
            if (epi == EPI_F32C) return launch_linear_t<float, ACX_WAVE2, EPI_F32C>(a, s);
        }
        return set_error(ACX_EUNSUPPORTED, "fp32 linear kernel supports the compensated stage only");
    }
    if (sys == ACX_ACOUSTIC4) {
        switch (epi) {
        case EPI_RHS: return launch_linear_t<double, ACX_ACOUSTIC4, EPI_RHS>(a, s);
        case EPI_AXPY: return launch_linear_t<double, ACX_ACOUSTIC4, EPI_AXPY>(a, s);
        case EPI_COMB: return launch_linear_t<double, ACX_ACOUSTIC4, EPI_COMB>(a, s);
        }
    } else {
        switch (epi) {
        case EPI_RHS: return launch_linear_t<double, ACX_WAVE2, EPI_RHS>(a, s);
        case EPI_AXPY: return launch_linear_t<double, ACX_WAVE2, EPI_AXPY>(a, s);
        case EPI_COMB: return launch_linear_t<double, ACX_WAVE2, EPI_COMB>(a, s);
        }
    }
    return set_error(ACX_EUNSUPPORTED, "linear kernel: unsupported epilogue");
}

} // namespace acx
