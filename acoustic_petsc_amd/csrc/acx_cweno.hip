// acx_cweno.hip -- fused RHS + SSP-stage kernel for the general (nonlinear)
// CWENO4 scheme: smoothness indicators, nonlinear weights, face averages,
// Riemann flux and divergence in one pass.
//
// Reconstruction: 3d/weno4/plane_wave/reconstruction.c:38-259 with the linear
// weights of the caller (the nonlinear setting of the reference is
// gamma_hi = gamma_lo = 0.85, Viscous-Multiphase-Fourth-Order/reconstruction.c:42-43).
// Only what the face averages need is formed: the 2x2 Gauss rule on a face
// annihilates every tangential basis term, so the face average of the value
// is c0 +- c1/2 + c4/6 +- c10/20 along x (c2,c5,c11 along y; c3,c6,c12 along
// z) and of the normal gradient (c1 +- c4 + 3/5 c10)/h.  The smoothness
// indicators still use the full 33-point stencil.
//
// Fluxes: LLF of PDEConsFlux (3d/weno3/LinearisedEuler/pde.c:35-95) for
// ACOUSTIC4, the central flux of pde.c:12-34 for WAVE2; both are linear in
// the face states, so the flux of the face averages equals the reference's
// average of Gauss-point fluxes.
//
// Layout: a block covers BX x BY reconstruction columns, the outer ring being
// the face-neighbour halo of the (BX-2) x (BY-2) output tile, and marches a
// chunk of z-planes.  The block keeps the five planes m-2..m+2 of its
// stencil region in an LDS ring (the next plane is loaded one iteration
// ahead into registers and stored after the last read of the slot it
// replaces).  x-neighbour face states and fluxes move by lane shuffles
// inside a 32-lane row, y-neighbours through LDS, z-neighbours stay in
// registers.  Selective mode (per-cell flags) replaces the full CWENO by
// the linear 4th-order polynomial wherever the flag bit of a component is 0.
#include <algorithm>

#include "acx_device.h"
#include "acx_internal.h"
#include "acx_cweno_core.h"

namespace acx {

enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3 };

// Face quantities of one component from its axis coefficients:
// ACOUSTIC4 -> face-averaged values, WAVE2 -> h * face-averaged normal gradient.
template <typename T, int SYS>
__device__ __forceinline__ void face_q(T c0, const Axis3<T> &a, T (&fm)[3], T (&fp)[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (SYS == ACX_ACOUSTIC4) {
            const T even = c0 + a.c2[d] * T(1.0 / 6.0), odd = T(0.5) * a.c1[d] + a.c3[d] * T(1.0 / 20.0);
            fm[d] = even - odd;
            fp[d] = even + odd;
        } else {
            const T e = a.c1[d] + T(0.6) * a.c3[d];
            fm[d] = e - a.c2[d];
            fp[d] = e + a.c2[d];
        }
    }
}

// Flux across a face normal to axis d from left (+face) and right (-face) states.
template <typename T, int SYS, int NS>
__device__ __forceinline__ void face_flux(int d, const T (&L)[NS], const T (&R)[NS], T K, T irho, T s, T (&F)[NS])
{
    if constexpr (SYS == ACX_ACOUSTIC4) {
        // 1/2 (F(Q_L) + F(Q_R) - s (Q_R - Q_L)), F_p = K u_n, F_un = p / rho0
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] = -s * (R[v] - L[v]);
        F[0] += K * (L[1 + d] + R[1 + d]);
        F[1 + d] += irho * (L[0] + R[0]);
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] *= T(0.5);
    } else {
        F[0] = T(0.5) * (L[0] + R[0]);
    }
}

// Shared-memory plan of one block (dynamic LDS):
//   ring  [5][NS][BY+4][BX+4]  planes m-2 .. m+2 of the stencil region
//   fym   [NS][BY][BX]         -y face state of each reconstruction column
//   Fy    [NS][BY][BX]         flux through the +y face of each column
template <typename T, int NS, int BX, int BY>
struct CwLds {
    static constexpr int LX = BX + 4, LY = BY + 4;
    static constexpr int NR = 5;
    static constexpr size_t ring = (size_t)NR * NS * LY * LX;
    static constexpr size_t bytes = (ring + 2ull * NS * BY * BX) * sizeof(T);
};

// Waves per SIMD the register allocation aims for.  The selective patch
// marches a few hundred marked tiles of zc + 2 planes each, and the time of one
// tile's march is what counts: with 2 waves per SIMD (one block per CU, no
// spills) against 4 (two blocks per CU, at most 128 VGPRs, 60 spilled), a
// steep pulse's selective step took 1.146 -> 1.076 ms at 256^3 (232 marked
// tiles, one round either way) and the same 3.14 ms at 384^3 (508 tiles: two
// rounds against one).  The full-grid kernel keeps the registers too.
template <bool SEL>
struct CwWaves {
    static constexpr int min_per_simd = SEL ? 2 : 1;
};

// WALLS = false: every axis periodic (wrap-only index maps, no ghost signs).
template <typename T, int SYS, int BX, int BY, int EPI, bool SEL, bool WALLS>
__global__ void __launch_bounds__(BX *BY) __attribute__((amdgpu_waves_per_eu(CwWaves<SEL>::min_per_simd)))
k_cweno(CwArgs p)
{
    constexpr int NV = SYS == ACX_ACOUSTIC4 ? 4 : 2;
    constexpr int NS = SYS == ACX_ACOUSTIC4 ? 4 : 1;
    constexpr int NT = BX * BY;
    using L = CwLds<T, NS, BX, BY>;
    constexpr int LX = L::LX, LY = L::LY;
    static_assert(BX == 32, "x exchange uses 32-lane row shuffles");
    extern __shared__ __align__(16) unsigned char smem[];
    T(*ring)[NS][LY][LX] = reinterpret_cast<T(*)[NS][LY][LX]>(smem);
    T(*s_fym)[BY][BX] = reinterpret_cast<T(*)[BY][BX]>(smem + L::ring * sizeof(T));
    T(*s_Fy)[BY][BX] = reinterpret_cast<T(*)[BY][BX]>(smem + (L::ring + (size_t)NS * BY * BX) * sizeof(T));

    const T *__restrict__ Q = static_cast<const T *>(p.q);
    const T *__restrict__ D = static_cast<const T *>(p.d);
    const T *__restrict__ U0 = static_cast<const T *>(p.u0);
    const T *__restrict__ LO = static_cast<const T *>(p.lo);
    T *__restrict__ OUT = static_cast<T *>(p.out);
    T *__restrict__ OUT2 = static_cast<T *>(p.out2);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    // Selective patch: the list of marked tiles (count, then tile indices in
    // ascending order), or every tile without a list; block b takes entries
    // b, b + grid, ...  The marked tiles cluster in space (along a front), and
    // consecutive entries go to consecutive blocks, i.e. round-robin over the
    // XCDs.  Only marked tiles are visited: with one block per tile of the
    // grid, the blocks that returned at once on unmarked tiles cost about
    // 16 us per stage at 256^3 and 21 us at 384^3 (4-plane tiles, measured).
    // Otherwise one tile per block, XCD-contiguous.
    int item = blockIdx.x;
    const int nitems = SEL ? (p.tlist ? p.tlist[0] : p.ntx * p.nty * p.ntz) : 1;
    if (SEL && item >= nitems) return;
    bool nfbad = false;   // non-finite watch of this thread's outputs
    do {
        const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
        const int tile = SEL ? (p.tlist ? p.tlist[1 + item] : item) : xcd_logical_tile(blockIdx.x, gridDim.x);
        const int gx0 = (tile % p.ntx) * (BX - 2) - 1, gy0 = ((tile / p.ntx) % p.nty) * (BY - 2) - 1;
        const int gx = gx0 + tx, gy = gy0 + ty;
        const int zb = (tile / (p.ntx * p.nty)) * p.zc, ze = min(zb + p.zc, nz);
        const bool interior = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2 && gx < nx && gy < ny;
        const size_t plane = (size_t)nx * ny;
        const int kx = WALLS ? p.bc[0] : 0, ky = WALLS ? p.bc[1] : 0, kz = WALLS ? p.bc[2] : 0;   // boundary kinds

        // own column (mapped) and its ghost sign
        float sx0 = 1.f, sy0 = 1.f;
        const uint32_t ocol = (uint32_t)(map_axis(gy, ny, ky, sy0) * nx + map_axis(gx, nx, kx, sx0));

        // ring cells this thread loads (at most 2 of the (BX+4)(BY+4) region)
        constexpr int NREG = LX * LY;
        constexpr int PER = (NREG + NT - 1) / NT;
        uint32_t rcol[PER];
        float rsx[PER], rsy[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int idx = tid + r * NT;
            rcol[r] = 0;
            rsx[r] = rsy[r] = 1.f;
            if (idx < NREG) {
                const int lx = idx % LX, ly = idx / LX;
                rcol[r] = (uint32_t)(map_axis(gy0 - 2 + ly, ny, ky, rsy[r]) * nx + map_axis(gx0 - 2 + lx, nx, kx, rsx[r]));
            }
        }
        const bool has_d = (EPI == EPI_F32C) && D != nullptr;
        auto zoff = [&](int z, float &s) -> size_t {
            s = 1.f;
            return (size_t)map_axis(z, nz, kz, s) * plane;
        };
        auto slot = [](int m) { return ((m % L::NR) + L::NR) % L::NR; };
        // Plane loads keep the raw words; the FP32 increment and the ghost sign
        // are applied when the plane is stored, so no wait sits at the load.
        struct Raw {
            Cell<T, NV> q, d;
            float s;
        };
        auto load_plane = [&](int z, Raw(&buf)[PER]) {
            float s;
            const size_t off = zoff(z, s);
#pragma unroll
            for (int r = 0; r < PER; ++r)
                if (tid + r * NT < NREG) {
                    buf[r].q = load_cell<T, NV>(Q, off + rcol[r]);
                    if (has_d) buf[r].d = load_cell<T, NV>(D, off + rcol[r]);
                    buf[r].s = s;
                }
        };
        auto store_plane = [&](int z, const Raw(&buf)[PER]) {
            const int sl = slot(z);
#pragma unroll
            for (int r = 0; r < PER; ++r) {
                const int idx = tid + r * NT;
                if (idx < NREG) {
                    Cell<T, NV> c = buf[r].q;
                    if (has_d) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) c.v[v] += buf[r].d.v[v];
                    }
                    if constexpr (WALLS) apply_bc_sign<T, NV>(c, rsx[r], rsy[r], buf[r].s);
#pragma unroll
                    for (int v = 0; v < NS; ++v) ring[sl][v][idx / LX][idx % LX] = c.v[v];
                }
            }
        };

        // prologue: planes zb-3 .. zb+1
        for (int z = zb - 3; z <= zb + 1; ++z) {
            Raw b0[PER];
            load_plane(z, b0);
            store_plane(z, b0);
        }
        __syncthreads();

        const T K = T(p.K), irho = T(p.invrho), sL = T(p.s);
        T zp_prev[NS], Fz_prev[NS], dxy_prev[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) { zp_prev[v] = T(0); Fz_prev[v] = T(0); dxy_prev[v] = T(0); }
        T vcen_prev = T(0);
        const int lx = tx + 2, ly = ty + 2;

        uint8_t fl_next = 0xff;
        // Selective patch: plane m+3 is requested in iteration m-1, right
        // after the plane before it went to LDS, instead of at the start of
        // iteration m.  (The compiler waits for these loads together with the
        // epilogue operands, before the results of iteration m-1 are stored,
        // so they overlap that iteration's fluxes.)  Measured with one block
        // per CU on a steep pulse at 256^3, in two runs beside a reference
        // build that took 1.15-1.16 ms in both: 1.135 -> 1.076 ms per step.
        Raw nxt_sel[PER];
        if constexpr (SEL) {
            float s;
            fl_next = p.flags[zoff(zb - 1, s) + ocol];
            load_plane(zb + 2, nxt_sel);
        }
        for (int m = zb - 1; m <= ze; ++m) {
            // prefetch plane m+3 of the region, the flags of plane m+1 and the
            // epilogue operands of plane m-1
            Raw nxt_full[PER];
            Raw(&nxt)[PER] = SEL ? nxt_sel : nxt_full;
            const bool pf = m + 3 <= ze + 2;
            if (!SEL && pf) load_plane(m + 3, nxt);
            const uint8_t fl = fl_next;
            if constexpr (SEL) {
                float s;
                if (m + 1 <= ze) fl_next = p.flags[zoff(m + 1, s) + ocol];
            }
            const bool fin = m >= zb + 1 && interior;
            const size_t cell = (size_t)(m - 1) * plane + ocol;
            Cell<T, NV> e0, e1, e2;
            if (fin) {
                // COMB: U0 is loaded at the epilogue instead (8 fewer live VGPRs
                // through the weights, where the registers spill)
                if constexpr (EPI == EPI_F32C) {
                    if (has_d) e0 = load_cell<T, NV>(D, cell);
                    if (p.final_stage) {
                        e1 = load_cell<T, NV>(LO, cell);
                        e2 = load_cell<T, NV>(Q, cell);
                    }
                }
            }
            T vcen = T(0);
            if constexpr (SYS == ACX_WAVE2) {
                float s;
                const size_t idx = (zoff(m, s) + ocol) * NV + 1;
                vcen = Q[idx];
                if constexpr (EPI == EPI_F32C) { if (D) vcen += D[idx]; }
            }

            const int sm2 = slot(m - 2), s0 = slot(m - 1), s1 = slot(m), s2 = slot(m + 1), sp2 = slot(m + 2);
            T fpx[NS], fmx[NS], fpy[NS], fmz[NS], fpz[NS];
#pragma unroll
            for (int v = 0; v < NS; ++v) {
                const T(*P0)[LX] = ring[s0][v];
                const T(*P1)[LX] = ring[s1][v];
                const T(*P2)[LX] = ring[s2][v];
                const T(*M2)[LX] = ring[sm2][v];
                const T(*Q2)[LX] = ring[sp2][v];
                // stencil slot q -> LDS (slot order of cweno_weights)
                auto get = [&](int q) -> T {
                    switch (q) {
                    case 0: return P1[ly][lx];
                    case 1: return P1[ly][lx - 2];
                    case 2: return P1[ly][lx - 1];
                    case 3: return P1[ly][lx + 1];
                    case 4: return P1[ly][lx + 2];
                    case 5: return P1[ly - 2][lx];
                    case 6: return P1[ly - 1][lx];
                    case 7: return P1[ly + 1][lx];
                    case 8: return P1[ly + 2][lx];
                    case 9: return M2[ly][lx];
                    case 10: return P0[ly][lx];
                    case 11: return P2[ly][lx];
                    case 12: return Q2[ly][lx];
                    case 13: return P1[ly + 1][lx + 1];   // xy: ++ +- -+ --
                    case 14: return P1[ly - 1][lx + 1];
                    case 15: return P1[ly + 1][lx - 1];
                    case 16: return P1[ly - 1][lx - 1];
                    case 17: return P2[ly + 1][lx];       // yz (y,z): ++ +- -+ --
                    case 18: return P0[ly + 1][lx];
                    case 19: return P2[ly - 1][lx];
                    case 20: return P0[ly - 1][lx];
                    case 21: return P2[ly][lx + 1];       // zx (z,x): ++ +- -+ --
                    case 22: return P2[ly][lx - 1];
                    case 23: return P0[ly][lx + 1];
                    case 24: return P0[ly][lx - 1];
                    case 25: return P2[ly + 1][lx + 1];   // corners (x,y,z): +++ -++ +-+ ++- --+ -+- +-- ---
                    case 26: return P2[ly + 1][lx - 1];
                    case 27: return P2[ly - 1][lx + 1];
                    case 28: return P0[ly + 1][lx + 1];
                    case 29: return P2[ly - 1][lx - 1];
                    case 30: return P0[ly + 1][lx - 1];
                    case 31: return P0[ly - 1][lx + 1];
                    default: return P0[ly - 1][lx - 1];
                    }
                };

                Axis3<T> ax;
                bool full = true;
                if constexpr (SEL) full = (fl >> v) & 1;
                if (full) {
                    cweno_axes_lds<T>(get, p, ax);
                } else {
                    axis_linear(get(1), get(2), get(0), get(3), get(4), ax.c1[0], ax.c2[0], ax.c3[0]);
                    axis_linear(get(5), get(6), get(0), get(7), get(8), ax.c1[1], ax.c2[1], ax.c3[1]);
                    axis_linear(get(9), get(10), get(0), get(11), get(12), ax.c1[2], ax.c2[2], ax.c3[2]);
                }
                T m3[3], p3[3];
                face_q<T, SYS>(get(0), ax, m3, p3);
                fmx[v] = m3[0]; fpx[v] = p3[0];
                s_fym[v][ty][tx] = m3[1];
                fpy[v] = p3[1];
                fmz[v] = m3[2]; fpz[v] = p3[2];
                asm volatile("" ::: "memory");   // one component's stencil live at a time
            }
            __syncthreads();   // B1: -y states visible; every read of plane m-1 is done

            if (pf) store_plane(m + 3, nxt);   // into the slot of plane m-2
            if (SEL && m + 4 <= ze + 2) load_plane(m + 4, nxt);

            T Fx[NS], Fy[NS];
            {
                T Lf[NS], Rf[NS];
#pragma unroll
                for (int v = 0; v < NS; ++v) { Lf[v] = fpx[v]; Rf[v] = __shfl_down(fmx[v], 1, 32); }
                face_flux<T, SYS, NS>(0, Lf, Rf, K, irho, sL, Fx);
#pragma unroll
                for (int v = 0; v < NS; ++v) { Lf[v] = fpy[v]; Rf[v] = ty + 1 < BY ? s_fym[v][ty + 1][tx] : T(0); }
                face_flux<T, SYS, NS>(1, Lf, Rf, K, irho, sL, Fy);
            }
#pragma unroll
            for (int v = 0; v < NS; ++v) s_Fy[v][ty][tx] = Fy[v];
            __syncthreads();   // B2: fluxes and plane m+2 visible

            T dxy[NS];
#pragma unroll
            for (int v = 0; v < NS; ++v) {
                const T Fxm = __shfl_up(Fx[v], 1, 32);
                const T Fym = ty > 0 ? s_Fy[v][ty - 1][tx] : T(0);
                dxy[v] = (Fx[v] - Fxm) + (Fy[v] - Fym);
            }

            if (m >= zb) {
                T Fz[NS], Lf[NS], Rf[NS];
#pragma unroll
                for (int v = 0; v < NS; ++v) { Lf[v] = zp_prev[v]; Rf[v] = fmz[v]; }
                face_flux<T, SYS, NS>(2, Lf, Rf, K, irho, sL, Fz);
                if (fin) {
                    Cell<T, NV> r;
                    if constexpr (SYS == ACX_ACOUSTIC4) {
                        const T ih = T(p.inv_h);
#pragma unroll
                        for (int v = 0; v < 4; ++v) r.v[v] = -ih * (dxy_prev[v] + (Fz[v] - Fz_prev[v]));
                    } else {
                        r.v[0] = vcen_prev;
                        r.v[1] = T(p.c2) * (dxy_prev[0] + (Fz[0] - Fz_prev[0]));
                    }
                    Cell<T, NV> o;
                    if constexpr (EPI == EPI_RHS) {
                        o = r;
                    } else if constexpr (EPI == EPI_AXPY || EPI == EPI_COMB) {
                        // Q of the own (in-domain) cell of plane m-1 is still in the ring
                        Cell<T, NV> q;
                        if constexpr (NS == NV) {
#pragma unroll
                            for (int v = 0; v < NV; ++v) q.v[v] = ring[s0][v][ly][lx];
                        } else {
                            q = load_cell<T, NV>(Q, cell);
                        }
                        if constexpr (EPI == EPI_AXPY) {
#pragma unroll
                            for (int v = 0; v < NV; ++v) o.v[v] = q.v[v] + T(p.c) * r.v[v];
                        } else {
                            e0 = load_cell<T, NV>(U0, cell);
#pragma unroll
                            for (int v = 0; v < NV; ++v) o.v[v] = T(p.a) * e0.v[v] + T(p.b) * q.v[v] + T(p.c) * r.v[v];
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) o.v[v] = (has_d ? T(p.b) * e0.v[v] : T(0)) + T(p.c) * r.v[v];
                        if (p.final_stage) {
                            Cell<T, NV> nh, nl;
#pragma unroll
                            for (int v = 0; v < NV; ++v) {
                                float t = (float)e1.v[v] + (float)o.v[v];
                                float sm, er;
                                two_sum((float)e2.v[v], t, sm, er);
                                nh.v[v] = sm;
                                nl.v[v] = er;
                            }
                            store_cell<T, NV>(OUT2, cell, nl);
                            o = nh;
                        }
                    }
                    if (p.nf) finite_acc<T, NV>(nfbad, o);
                    store_cell<T, NV>(OUT, cell, o);
                }
#pragma unroll
                for (int v = 0; v < NS; ++v) Fz_prev[v] = Fz[v];
            }
#pragma unroll
            for (int v = 0; v < NS; ++v) { zp_prev[v] = fpz[v]; dxy_prev[v] = dxy[v]; }
            vcen_prev = vcen;
        }
        if constexpr (SEL) {
            __syncthreads();   // the next tile's prologue rewrites the ring
            item += gridDim.x;
        }
    } while (SEL && item < nitems);
    watch_flush(p.nf, nfbad);
}

// ---------------------------------------------------------------------------

constexpr int CW_BX = 32, CW_BY = 16;   // reconstruction columns per block (output tile 30 x 14)

// One block per column of up to TM_ZT CWENO tiles along z: a tile is marked
// when a cell of it, or a face neighbour of one (a cell of the one-cell ring
// outside the tile in exactly one coordinate, periodic wrap), is flagged;
// their outputs use the flagged reconstruction.  Every entry is written (0 or
// 1), one store per tile.
constexpr int TM_ZT = 8;

__global__ void __launch_bounds__(256) k_tile_mark(const uint8_t *__restrict__ flags, int nx, int ny, int nz, int zc,
                                                   int ntx, int nty, int ntz, uint8_t *__restrict__ mask)
{
    static_assert(CW_BX - 2 + 2 <= 32 && CW_BY - 2 + 2 <= 16, "one thread column per box column, two row passes");
    const int col = blockIdx.x % (ntx * nty), tz0 = (blockIdx.x / (ntx * nty)) * TM_ZT;
    const int x0 = (col % ntx) * (CW_BX - 2), y0 = (col / ntx) * (CW_BY - 2);
    const int bx = min(CW_BX - 2, nx - x0) + 2, by = min(CW_BY - 2, ny - y0) + 2;
    const int lx = threadIdx.x & 31;
    int x = x0 - 1 + lx;
    x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
    const int ex = (lx == 0 || lx == bx - 1);
    for (int tz = tz0; tz < min(tz0 + TM_ZT, ntz); ++tz) {
        const int z0 = tz * zc, bz = min(zc, nz - z0) + 2;
        int any = 0;
        if (lx < bx) {
            for (int ly = threadIdx.x >> 5; ly < by; ly += 8) {
                int y = y0 - 1 + ly;
                y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
                const int exy = ex + (ly == 0 || ly == by - 1);
                if (exy > 1) continue;
                for (int lz = 0; lz < bz; ++lz) {
                    if (exy + (lz == 0 || lz == bz - 1) > 1) continue;
                    int z = z0 - 1 + lz;
                    z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
                    any |= flags[((size_t)z * ny + y) * nx + x];
                }
            }
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) mask[((size_t)tz * nty) * ntx + col] = any ? 1 : 0;
    }
}

// The marked tiles as a list: list[0] = their number, list[1 ..] their
// indices in ascending order.  One block walks the mask in chunks of 1024
// entries; a chunk's marked entries are placed by a wavefront ballot and the
// counts of the wavefronts before it.
constexpr int TL_NT = 1024;

__global__ void __launch_bounds__(TL_NT) k_tile_list(const uint8_t *__restrict__ mask, int tiles, int *__restrict__ list)
{
    __shared__ int wcount[TL_NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int base = 0;
    for (int c0 = 0; c0 < tiles; c0 += TL_NT) {
        const int t = c0 + (int)threadIdx.x;
        const bool on = t < tiles && mask[t] != 0;
        const unsigned long long b = __ballot(on);
        if (lane == 0) wcount[w] = __popcll(b);
        __syncthreads();
        int before = base, total = 0;
#pragma unroll
        for (int k = 0; k < TL_NT / 64; ++k) {
            before += k < w ? wcount[k] : 0;
            total += wcount[k];
        }
        if (on) list[1 + before + __popcll(b & ((1ull << lane) - 1ull))] = t;
        base += total;
        __syncthreads();   // wcount is rewritten by the next chunk
    }
    if (threadIdx.x == 0) list[0] = base;
}

size_t cweno_sel_tiles(const CwArgs &a)
{
    const size_t ntx = (a.nx + CW_BX - 3) / (CW_BX - 2), nty = (a.ny + CW_BY - 3) / (CW_BY - 2);
    return ntx * nty * ((a.nz + a.zc - 1) / a.zc);
}

int launch_cweno_tilemask(const CwArgs &a, const uint8_t *flags, uint8_t *mask, hipStream_t s)
{
    const int ntx = (a.nx + CW_BX - 3) / (CW_BX - 2), nty = (a.ny + CW_BY - 3) / (CW_BY - 2);
    const int ntz = (a.nz + a.zc - 1) / a.zc;
    const size_t tiles = cweno_sel_tiles(a);
    if (tiles > (size_t)INT32_MAX - 1) return set_error(ACX_EUNSUPPORTED, "tile mask: %zu patch tiles", tiles);
    hipLaunchKernelGGL(k_tile_mark, dim3((unsigned)((size_t)ntx * nty * ((ntz + TM_ZT - 1) / TM_ZT))), dim3(256), 0, s,
                       flags, a.nx, a.ny, a.nz, a.zc, ntx, nty, ntz, mask);
    int rc = hip_check(hipGetLastError(), "tile mask launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_tile_list, dim3(1), dim3(TL_NT), 0, s, mask, (int)tiles,
                       reinterpret_cast<int *>(mask + sel_list_offset(tiles)));
    return hip_check(hipGetLastError(), "tile list launch");
}

template <typename T, int SYS, int EPI, bool SEL, bool WALLS>
static int launch_cw_tw(const CwArgs &a, hipStream_t s)
{
    constexpr int BX = CW_BX, BY = CW_BY;
    CwArgs b = a;
    b.ntx = (a.nx + BX - 3) / (BX - 2);
    b.nty = (a.ny + BY - 3) / (BY - 2);
    if (b.zc <= 0) {
        // round model with a start of 4 planes (the recomputed ring planes and
        // the prologue).  Measured (FP64, gamma = 0.85): 64 planes at 256^3
        // (+3.7 % over 32; 96 leaves a one-workgroup last round), 96 at 384^3
        // (+2-4 % over 64; 128 -12 %).
        static int per_cu = 0;
        if (!per_cu) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, reinterpret_cast<const void *>(&k_cweno<T, SYS, BX, BY, EPI, SEL, WALLS>), BX * BY,
                    CwLds<T, SYS == ACX_ACOUSTIC4 ? 4 : 1, BX, BY>::bytes) != hipSuccess || n <= 0)
                n = 1;
            per_cu = n;
        }
        b.zc = auto_zc(b.ntx, b.nty, a.nz, device_cus() * per_cu, 4);
    }
    b.ntz = (a.nz + b.zc - 1) / b.zc;
    size_t blocks = (size_t)b.ntx * b.nty * b.ntz;
    if (SEL && b.tlist) {
        // the marked tiles of the list, at most one round of resident blocks
        static int per_cu_sel = 0;
        if (!per_cu_sel) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, reinterpret_cast<const void *>(&k_cweno<T, SYS, BX, BY, EPI, SEL, WALLS>), BX * BY,
                    CwLds<T, SYS == ACX_ACOUSTIC4 ? 4 : 1, BX, BY>::bytes) != hipSuccess || n <= 0)
                n = 1;
            per_cu_sel = n;
        }
        blocks = std::min(blocks, (size_t)device_cus() * per_cu_sel);
    }
    constexpr int NS = SYS == ACX_ACOUSTIC4 ? 4 : 1;
    constexpr size_t lds = CwLds<T, NS, BX, BY>::bytes;
    static bool attr_set = false;
    if (!attr_set) {
        int rc = hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_cweno<T, SYS, BX, BY, EPI, SEL, WALLS>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                           "k_cweno LDS attribute");
        if (rc) return rc;
        attr_set = true;
    }
    dim3 blk(BX, BY, 1);
    dim3 grd((unsigned)blocks, 1, 1);
    hipLaunchKernelGGL((k_cweno<T, SYS, BX, BY, EPI, SEL, WALLS>), grd, blk, lds, s, b);
    return hip_check(hipGetLastError(), "k_cweno launch");
}

template <typename T, int SYS, int EPI, bool SEL>
static int launch_cw_t(const CwArgs &a, hipStream_t s)
{
    if (a.bc[0] != ACX_PERIODIC || a.bc[1] != ACX_PERIODIC || a.bc[2] != ACX_PERIODIC)
        return launch_cw_tw<T, SYS, EPI, SEL, true>(a, s);
    return launch_cw_tw<T, SYS, EPI, SEL, false>(a, s);
}

int launch_cweno(const CwArgs &a, int sys, int epi, bool fp32, bool selective, hipStream_t s)
{
    if (fp32) {
        if (epi != EPI_F32C || sys != ACX_ACOUSTIC4)
            return set_error(ACX_EUNSUPPORTED, "fp32 cweno kernel: compensated ACOUSTIC4 stage only");
        return selective ? launch_cw_t<float, ACX_ACOUSTIC4, EPI_F32C, true>(a, s)
                         : launch_cw_t<float, ACX_ACOUSTIC4, EPI_F32C, false>(a, s);
    }
    if (selective) return set_error(ACX_EUNSUPPORTED, "fp64 cweno kernel: selective mode not built");
    if (sys == ACX_ACOUSTIC4) {
        switch (epi) {
        case EPI_RHS: return launch_cw_t<double, ACX_ACOUSTIC4, EPI_RHS, false>(a, s);
        case EPI_AXPY: return launch_cw_t<double, ACX_ACOUSTIC4, EPI_AXPY, false>(a, s);
        case EPI_COMB: return launch_cw_t<double, ACX_ACOUSTIC4, EPI_COMB, false>(a, s);
        }
    } else {
        switch (epi) {
        case EPI_RHS: return launch_cw_t<double, ACX_WAVE2, EPI_RHS, false>(a, s);
        case EPI_AXPY: return launch_cw_t<double, ACX_WAVE2, EPI_AXPY, false>(a, s);
        case EPI_COMB: return launch_cw_t<double, ACX_WAVE2, EPI_COMB, false>(a, s);
        }
    }
    return set_error(ACX_EUNSUPPORTED, "cweno kernel: unsupported epilogue");
}

} // namespace acx
