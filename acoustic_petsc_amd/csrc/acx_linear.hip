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
// chunk of z-planes.  Each plane is read once: staged in LDS with an x/y
// halo of 3 (double-buffered, one barrier per plane) for its in-plane
// stencils, and pushed into a per-thread register queue that supplies the
// z-stencil three planes later.  HBM traffic per stage is one read of Q
// (+ U0) and one write of OUT; halo rows are read by neighbouring tiles at
// the same plane index, so they meet in L2 (tiles are ordered per XCD).
#include "acx_linear_core.h"

namespace acx {

// PF = 0: centre-plane march (PF = -1: with late epilogue loads; PF = -2:
// with the 5-plane halo ring); PF >= 1: lead-plane march with loads PF
// planes ahead.  WALLS = false is the
// all-periodic instantiation (no ghost signs).
template <typename T, int SYS, int BX, int BY, int EPI, int PF, int WPE, bool WALLS, bool NT = false>
__global__ void __launch_bounds__(BX *BY) __attribute__((amdgpu_waves_per_eu(WPE)))
k_linear(LinArgs p)
{
    using Tile = LinTile<T, SYS, BX, BY>;
    __shared__ T lds[PF == -2 ? 5 : 2][Tile::NS][BY + 6][BX + 6];
    const int tile = xcd_logical_tile(blockIdx.x, gridDim.x);
    const int x0 = (tile % p.ntx) * BX, y0 = ((tile / p.ntx) % p.nty) * BY, zb = (tile / (p.ntx * p.nty)) * p.zc;
    if constexpr (PF <= 0)
        linear_tile_centre<T, SYS, BX, BY, EPI, (PF == -1), WALLS, NT>(p, x0, y0, zb, lds);
    else
        linear_tile<T, SYS, BX, BY, EPI, PF, WALLS, NT>(p, x0, y0, zb, lds);
}

// LDS-DMA march (acx_linear_core.h): FP64 ACOUSTIC4 RHS / AXPY stages on
// all-periodic grids; one block of 512 threads per CU (the ring is 160.5 KB).
// RS: row-complete result stores, taken when the launch needs more than one
// round of workgroups.
template <int EPI, bool RS>
__global__ void __launch_bounds__(DmaPlan::BX *DmaPlan::BY) k_linear_dma(LinArgs p)
{
    __shared__ __attribute__((aligned(16))) unsigned char ring[DmaPlan::NSLOT * DmaPlan::SLOT];
    const int tile = xcd_logical_tile(blockIdx.x, gridDim.x);
    const int x0 = (tile % p.ntx) * DmaPlan::BX, y0 = ((tile / p.ntx) % p.nty) * DmaPlan::BY;
    const int zb = (tile / (p.ntx * p.nty)) * p.zc;
    linear_tile_dma<EPI, RS>(p, x0, y0, zb, ring);
}

template <int EPI>
static int launch_linear_dma(const LinArgs &a, hipStream_t s)
{
    LinArgs b = a;
    b.ntx = (a.nx + DmaPlan::BX - 1) / DmaPlan::BX;
    b.nty = (a.ny + DmaPlan::BY - 1) / DmaPlan::BY;
    static int per_cu = 0;
    if (!per_cu) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_linear_dma<EPI, false>),
                                                         DmaPlan::BX * DmaPlan::BY, 0) != hipSuccess || n <= 0)
            n = 1;
        per_cu = n;
    }
    const int slots = device_cus() * per_cu;
    if (b.zc <= 0) b.zc = auto_zc(b.ntx, b.nty, a.nz, slots, 10);
    b.ntz = (a.nz + b.zc - 1) / b.zc;
    const long long wgs = (long long)b.ntx * b.nty * b.ntz;
    if (wgs > slots)
        hipLaunchKernelGGL((k_linear_dma<EPI, true>), dim3((unsigned)wgs), dim3(DmaPlan::BX, DmaPlan::BY), 0, s, b);
    else
        hipLaunchKernelGGL((k_linear_dma<EPI, false>), dim3((unsigned)wgs), dim3(DmaPlan::BX, DmaPlan::BY), 0, s, b);
    return hip_check(hipGetLastError(), "k_linear_dma launch");
}

// WAVE2 RHS / AXPY stages on the same march: 80 KB of LDS, two blocks per CU
// (registers capped for 4 waves per SIMD).
template <int EPI>
__global__ void __launch_bounds__(DmaPlan::BX *DmaPlan::BY) __attribute__((amdgpu_waves_per_eu(4)))
k_linear_dma_w2(LinArgs p)
{
    __shared__ __attribute__((aligned(16))) unsigned char ring[DmaPlanW2::NSLOT * DmaPlanW2::SLOT];
    const int tile = xcd_logical_tile(blockIdx.x, gridDim.x);
    const int x0 = (tile % p.ntx) * DmaPlan::BX, y0 = ((tile / p.ntx) % p.nty) * DmaPlan::BY;
    const int zb = (tile / (p.ntx * p.nty)) * p.zc;
    linear_tile_dma_w2<EPI>(p, x0, y0, zb, ring);
}

template <int EPI>
static int launch_linear_dma_w2(const LinArgs &a, hipStream_t s)
{
    LinArgs b = a;
    b.ntx = (a.nx + DmaPlan::BX - 1) / DmaPlan::BX;
    b.nty = (a.ny + DmaPlan::BY - 1) / DmaPlan::BY;
    static int per_cu = 0;
    if (!per_cu) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_linear_dma_w2<EPI>),
                                                         DmaPlan::BX * DmaPlan::BY, 0) != hipSuccess || n <= 0)
            n = 1;
        per_cu = n;
    }
    if (b.zc <= 0) b.zc = auto_zc(b.ntx, b.nty, a.nz, device_cus() * per_cu, 10);
    b.ntz = (a.nz + b.zc - 1) / b.zc;
    hipLaunchKernelGGL(k_linear_dma_w2<EPI>, dim3((unsigned)(b.ntx * b.nty * b.ntz)), dim3(DmaPlan::BX, DmaPlan::BY),
                       0, s, b);
    return hip_check(hipGetLastError(), "k_linear_dma_w2 launch");
}

// ---------------------------------------------------------------------------

// Automatic z-chunk (auto_zc, acx_api.hip): measured for these kernels (ACOUSTIC4,
// 256^3 and 384^3), the round model ranks the z-chunks as the timings do for
// the FP64 halo-ring kernel (one workgroup per CU: 128 planes at 256^3, +4 %
// over 64; 64 at 384^3) and the FP32 lead-plane kernel (two per CU: 64 at
// 256^3, 128 at 384^3, +3 % over 64).
template <typename T, int SYS, int EPI, int PF, int WPE, int BX, int BY, bool WALLS = false, bool NT = false>
static int launch_linear_v(const LinArgs &a, hipStream_t s)
{
    LinArgs b = a;
    b.ntx = (a.nx + BX - 1) / BX;
    b.nty = (a.ny + BY - 1) / BY;
    if (b.zc <= 0) {
        // workgroups of this instantiation that fit a CU at once (queried once)
        static int per_cu = 0;
        if (!per_cu) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, reinterpret_cast<const void *>(&k_linear<T, SYS, BX, BY, EPI, PF, WPE, WALLS, NT>), BX * BY, 0) !=
                    hipSuccess || n <= 0)
                n = 1;
            per_cu = n;
        }
        // At most two blocks per CU count as slots: the round model takes the
        // smallest z-chunk that fills one round, and with four co-resident
        // blocks of the WAVE2 halo-ring kernel that was 32 planes, whose
        // z-halo re-reads cost more than the round saved (measured at 256^3:
        // WAVE2 step 0.458 ms with the COMB stage at 32 planes, 0.450 at 64).
        b.zc = auto_zc(b.ntx, b.nty, a.nz, device_cus() * (per_cu < 2 ? per_cu : 2), 10);
    }
    b.ntz = (a.nz + b.zc - 1) / b.zc;
    dim3 blk(BX, BY, 1);
    dim3 grd(b.ntx * b.nty * b.ntz, 1, 1);
    hipLaunchKernelGGL((k_linear<T, SYS, BX, BY, EPI, PF, WPE, WALLS, NT>), grd, blk, 0, s, b);
    return hip_check(hipGetLastError(), "k_linear launch");
}

// One launch shape per precision, chosen by timing every candidate at 256^3
// and 384^3 on MI355X: FP64, the centre-plane march with the 5-plane halo
// ring on 32 x 16 tiles, nontemporal epilogue loads and stores; FP32, the
// lead-plane march on 64 x 8 tiles with loads one plane ahead and
// nontemporal stores.  An all-periodic grid takes the instantiation without
// ghost-sign work, a grid with a wall on any axis the one with it.
template <typename T, int SYS, int EPI>
static int launch_linear_t(const LinArgs &a, hipStream_t s)
{
    const bool walls = a.bc[0] != ACX_PERIODIC || a.bc[1] != ACX_PERIODIC || a.bc[2] != ACX_PERIODIC;
    if constexpr (sizeof(T) == 8 && SYS == ACX_ACOUSTIC4 && (EPI == EPI_RHS || EPI == EPI_AXPY)) {
        // the LDS-DMA march (measured at 256^3: AXPY stage 207 -> 196 us,
        // step 0.958 -> 0.905 ms): periodic grids whose plane offsets fit 32 bits
        if (!walls && (size_t)a.nx * a.ny * 4 < 0x7fffffffu)
            return launch_linear_dma<EPI>(a, s);
    }
    if constexpr (sizeof(T) == 8 && SYS == ACX_WAVE2 && (EPI == EPI_RHS || EPI == EPI_AXPY)) {
        // measured: the WAVE2 step 0.485 -> 0.459 ms at 256^3, 1.904 -> 1.760 ms
        // at 384^3, bit-identical to the halo-ring kernel
        if (!walls && (size_t)a.nx * a.ny * 2 < 0x7fffffffu)
            return launch_linear_dma_w2<EPI>(a, s);
    }
    if constexpr (sizeof(T) == 8) {
        return walls ? launch_linear_v<T, SYS, EPI, -2, 1, 32, 16, true, true>(a, s)
                     : launch_linear_v<T, SYS, EPI, -2, 1, 32, 16, false, true>(a, s);
    } else {
        return walls ? launch_linear_v<T, SYS, EPI, 1, 1, 64, 8, true, true>(a, s)
                     : launch_linear_v<T, SYS, EPI, 1, 1, 64, 8, false, true>(a, s);
    }
}

int launch_linear(const LinArgs &a, int sys, int epi, bool fp32, hipStream_t s)
{
    if (fp32) {
        // the stage that ends the step has its own instantiation (EPI_F32CF)
        if (sys == ACX_ACOUSTIC4) {
            if (epi == EPI_F32C)
                return a.final_stage ? launch_linear_t<float, ACX_ACOUSTIC4, EPI_F32CF>(a, s)
                                     : launch_linear_t<float, ACX_ACOUSTIC4, EPI_F32C>(a, s);
        } else {
            if (epi == EPI_F32C)
                return a.final_stage ? launch_linear_t<float, ACX_WAVE2, EPI_F32CF>(a, s)
                                     : launch_linear_t<float, ACX_WAVE2, EPI_F32C>(a, s);
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
