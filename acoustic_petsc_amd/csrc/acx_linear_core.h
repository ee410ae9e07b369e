// acx_linear_core.h -- one tile of the fused linear-scheme stage (see
// acx_linear.hip for the method); shared by the per-grid kernel and the
// persistent ensemble kernel.
#pragma once
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
// EPI_F32CF: the compensated FP32 stage that ends the step (writes the new
// HI and LO; p.final_stage is implied), instantiated on its own so that the
// lead-plane march can keep HI of the last four planes in registers for the
// epilogue's exact sum instead of loading it again three planes later.
enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3, EPI_F32CF = 4 };

// One tile: columns [x0, x0+BX) x [y0, y0+BY), planes [zb, ze) of the grid
// described by p (shared by the per-grid kernel and the ensemble kernel).
// WALLS = false: every axis periodic (no ghost signs, wrap-only index maps).
// NT: results stored nontemporally (streamed past L2).
template <typename T, int SYS, int BX, int BY, int EPI, int PF = 2, bool WALLS = true, bool NT = false>
__device__ __forceinline__ void linear_tile(const LinArgs &p, int x0, int y0, int zb,
                                            T (&lds)[2][LinTile<T, SYS, BX, BY>::NS][BY + 6][BX + 6])
{
    using Tile = LinTile<T, SYS, BX, BY>;
    constexpr int NV = Tile::NV, NS = Tile::NS;
    static_assert(BX * BY >= Tile::NHALO, "every halo cell of a plane needs a thread of its own");

    const T *__restrict__ Q = static_cast<const T *>(p.q);
    const T *__restrict__ D = static_cast<const T *>(p.d);       // F32C: increment added to Q
    const T *__restrict__ U0 = static_cast<const T *>(p.u0);
    const T *__restrict__ LO = static_cast<const T *>(p.lo);
    T *__restrict__ OUT = static_cast<T *>(p.out);
    T *__restrict__ OUT2 = static_cast<T *>(p.out2);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int ze = min(zb + p.zc, nz);
    const bool active = (x0 + tx < nx) && (y0 + ty < ny);
    const size_t plane = (size_t)nx * ny;
    const int bcx = WALLS ? p.bc[0] : 0, bcy = WALLS ? p.bc[1] : 0, bcz = WALLS ? p.bc[2] : 0;

    // own column
    float sgx = 1.f, sgy = 1.f;
    const int mx = map_axis(x0 + tx, nx, bcx, sgx);
    const int my = map_axis(y0 + ty, ny, bcy, sgy);
    const size_t col = (size_t)my * nx + mx;

    // halo cell owned by this thread (x halo for rows 0..BY-1, y halo for cols 0..BX-1)
    int hx = -1, hy = -1;
    size_t hcol = 0;
    float hsx = 1.f, hsy = 1.f;
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
        int gx = map_axis(x0 + hx - 3, nx, bcx, hsx);
        int gy = map_axis(y0 + hy - 3, ny, bcy, hsy);
        hcol = (size_t)gy * nx + gx;
    }

    auto zmap = [&](int z, float &s) -> size_t {
        s = 1.f;
        return (size_t)map_axis(z, nz, bcz, s) * plane;
    };

    // Lead-plane march.  At iteration j the plane j is staged in LDS (own
    // column and halo loaded for the same plane index as the neighbouring
    // blocks load those rows as their interior, so shared rows meet in L2),
    // its in-plane (x, y) part of the RHS is formed and queued, and plane
    // c = j - 3, whose z-neighbourhood c-3..c+3 is now complete in registers,
    // is finished and stored.
    //
    // Loads run PF iterations ahead.  A slot holds the raw loaded words (the
    // ghost sign and the FP32 increment are applied when the slot is
    // consumed, so nothing waits on a load at issue time), and the loop is
    // unrolled PF times so every slot keeps fixed registers while in flight.
    constexpr int NP = SYS == ACX_ACOUSTIC4 ? 4 : 1;
    const T inv48h = T(p.inv48h), Kc = T(p.K), irho = T(p.invrho), sLLF = T(p.s);
    const T lap = T(p.lapc);   // WAVE2: c^2 / (48 h^2)
    const int jbeg = zb - 3, jend = ze + 2;
    constexpr bool F32C = EPI == EPI_F32C || EPI == EPI_F32CF;
    const bool has_d = F32C && D != nullptr;
    const bool fin_stage = EPI == EPI_F32CF || (EPI == EPI_F32C && p.final_stage);

    struct Slot {
        Cell<T, NV> q, d, hq, hd, e0, e1, e2;
        float zs;
    };
    Slot sl[PF];

    auto issue = [&](Slot &S, int j) {
        if (j > jend) return;
        const size_t off = zmap(j, S.zs);
        S.q = load_cell<T, NV>(Q, off + col);
        if (has_d) S.d = load_cell<T, NV>(D, off + col);
        if (hx >= 0 && j >= zb && j < ze) {
            S.hq = load_cell<T, NV>(Q, off + hcol);
            if (has_d) S.hd = load_cell<T, NV>(D, off + hcol);
        }
        const int c = j - 3;
        if (c >= zb && c < ze && active) {
            const size_t cidx = (size_t)c * plane + col;
            if constexpr (EPI == EPI_COMB) S.e0 = load_cell<T, NV>(U0, cidx);
            if constexpr (F32C) {
                if (fin_stage) {
                    // LO is read once: streamed past L2 with the nontemporal stores
                    if constexpr (NT) S.e1 = load_cell_nt<T, NV>(LO, cidx);
                    else S.e1 = load_cell<T, NV>(LO, cidx);
                    // HI of plane c, exact (EPI_F32CF: from the register queue)
                    if constexpr (EPI == EPI_F32C) S.e2 = load_cell<T, NV>(Q, cidx);
                }
            }
        }
    };

    Cell<T, NV> qz[7];   // column values of planes j-6 .. j after the push
    T xyq[4][NP];        // in-plane RHS parts of planes j-3 .. j
    // F32C: the increment D of planes j-3 .. j, kept from the stencil loads for
    // the epilogue of plane j-3 (a second load of it three planes later missed L2)
    Cell<T, NV> dq[F32C ? 4 : 1];
    // EPI_F32CF: HI of planes j-3 .. j as loaded (before adding D)
    Cell<T, NV> hq[EPI == EPI_F32CF ? 4 : 1];
#pragma unroll
    for (int k = 0; k < (F32C ? 4 : 1); ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) dq[k].v[v] = T(0);
#pragma unroll
    for (int k = 0; k < (EPI == EPI_F32CF ? 4 : 1); ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) hq[k].v[v] = T(0);
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) qz[k].v[v] = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int v = 0; v < NP; ++v) xyq[k][v] = T(0);

#pragma unroll
    for (int u = 0; u < PF; ++u) issue(sl[u], jbeg + u);

    bool nfbad = false;   // non-finite watch of this thread's outputs
    int buf = 0;
    for (int jb = jbeg; jb <= jend; jb += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = jb + u;
            if (j > jend) break;
            Slot &S = sl[u];
            const bool stage_xy = (j >= zb) && (j < ze);
            const int c = j - 3;
            const bool finish = (c >= zb) && (c < ze) && active;

            // consume the slot of plane j
            Cell<T, NV> cur = S.q;
            if constexpr (EPI == EPI_F32CF) {
#pragma unroll
                for (int k = 0; k < 3; ++k) hq[k] = hq[k + 1];
                hq[3] = S.q;
            }
            if (has_d) {
#pragma unroll
                for (int v = 0; v < NV; ++v) cur.v[v] += S.d.v[v];
                if constexpr (F32C) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) dq[k] = dq[k + 1];
                    dq[3] = S.d;
                }
            }
            if constexpr (WALLS) apply_bc_sign<T, NV>(cur, sgx, sgy, S.zs);

            T nxy[NP];
#pragma unroll
            for (int v = 0; v < NP; ++v) nxy[v] = T(0);
            if (stage_xy) {
#pragma unroll
                for (int v = 0; v < NS; ++v) lds[buf][v][ty + 3][tx + 3] = cur.v[v];
                if (hx >= 0) {
                    Cell<T, NV> hcur = S.hq;
                    if (has_d) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) hcur.v[v] += S.hd.v[v];
                    }
                    if constexpr (WALLS) apply_bc_sign<T, NV>(hcur, hsx, hsy, S.zs);
#pragma unroll
                    for (int v = 0; v < NS; ++v) lds[buf][v][hy][hx] = hcur.v[v];
                }
                __syncthreads();
                if constexpr (SYS == ACX_ACOUSTIC4) {
                    T accB[4], dAx = T(0), dAy = T(0), dAvel = T(0);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const T *row = &lds[buf][v][ty + 3][tx + 3];
                        const T q0 = cur.v[v];
                        const T xm3 = row[-3], xm2 = row[-2], xm1 = row[-1], xp1 = row[1], xp2 = row[2], xp3 = row[3];
                        const T ym3 = lds[buf][v][ty][tx + 3], ym2 = lds[buf][v][ty + 1][tx + 3];
                        const T ym1 = lds[buf][v][ty + 2][tx + 3];
                        const T yp1 = lds[buf][v][ty + 4][tx + 3], yp2 = lds[buf][v][ty + 5][tx + 3];
                        const T yp3 = lds[buf][v][ty + 6][tx + 3];
                        accB[v] = stencil_B7(xm3, xm2, xm1, q0, xp1, xp2, xp3) + stencil_B7(ym3, ym2, ym1, q0, yp1, yp2, yp3);
                        if (v == 0) {
                            dAx = stencil_A7(xm3, xm2, xm1, xp1, xp2, xp3);
                            dAy = stencil_A7(ym3, ym2, ym1, yp1, yp2, yp3);
                        } else if (v == 1) {
                            dAvel += stencil_A7(xm3, xm2, xm1, xp1, xp2, xp3);
                        } else if (v == 2) {
                            dAvel += stencil_A7(ym3, ym2, ym1, yp1, yp2, yp3);
                        }
                    }
                    nxy[0] = sLLF * accB[0] - Kc * dAvel;
                    nxy[1] = sLLF * accB[1] - irho * dAx;
                    nxy[2] = sLLF * accB[2] - irho * dAy;
                    nxy[3] = sLLF * accB[3];
                } else {
                    const T *row = &lds[buf][0][ty + 3][tx + 3];
                    const T q0 = cur.v[0];
                    nxy[0] = stencil_L7(row[-3], row[-2], row[-1], q0, row[1], row[2], row[3]) +
                             stencil_L7(lds[buf][0][ty][tx + 3], lds[buf][0][ty + 1][tx + 3], lds[buf][0][ty + 2][tx + 3],
                                        q0, lds[buf][0][ty + 4][tx + 3], lds[buf][0][ty + 5][tx + 3],
                                        lds[buf][0][ty + 6][tx + 3]);
                }
                buf ^= 1;
            }

            // push plane j
#pragma unroll
            for (int k = 0; k < 6; ++k) qz[k] = qz[k + 1];
            qz[6] = cur;
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int v = 0; v < NP; ++v) xyq[k][v] = xyq[k + 1][v];
#pragma unroll
            for (int v = 0; v < NP; ++v) xyq[3][v] = nxy[v];

            if (finish) {
                // z part of plane c from qz[0..6] = planes c-3..c+3
                const size_t cidx = (size_t)c * plane + col;
                Cell<T, NV> r;
                if constexpr (SYS == ACX_ACOUSTIC4) {
                    T bz[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        bz[v] = stencil_B7(qz[0].v[v], qz[1].v[v], qz[2].v[v], qz[3].v[v], qz[4].v[v], qz[5].v[v],
                                           qz[6].v[v]);
                    const T azp = stencil_A7(qz[0].v[0], qz[1].v[0], qz[2].v[0], qz[4].v[0], qz[5].v[0], qz[6].v[0]);
                    const T azw = stencil_A7(qz[0].v[3], qz[1].v[3], qz[2].v[3], qz[4].v[3], qz[5].v[3], qz[6].v[3]);
                    r.v[0] = inv48h * (xyq[0][0] + sLLF * bz[0] - Kc * azw);
                    r.v[1] = inv48h * (xyq[0][1] + sLLF * bz[1]);
                    r.v[2] = inv48h * (xyq[0][2] + sLLF * bz[2]);
                    r.v[3] = inv48h * (xyq[0][3] + sLLF * bz[3] - irho * azp);
                } else {
                    r.v[0] = qz[3].v[1];     // du/dt = v (rhs_function.c:330)
                    r.v[1] = lap * (xyq[0][0] + stencil_L7(qz[0].v[0], qz[1].v[0], qz[2].v[0], qz[3].v[0], qz[4].v[0],
                                                           qz[5].v[0], qz[6].v[0]));
                }
                Cell<T, NV> o;
                if constexpr (EPI == EPI_RHS) {
                    o = r;
                } else if constexpr (EPI == EPI_AXPY) {
                    const T cc = T(p.c);
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = qz[3].v[v] + cc * r.v[v];
                } else if constexpr (EPI == EPI_COMB) {
                    const T aa = T(p.a), bb = T(p.b), cc = T(p.c);
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = aa * S.e0.v[v] + bb * qz[3].v[v] + cc * r.v[v];
                } else {
                    // compensated FP32: increments relative to the step's base (HI + LO)
                    const T bb = T(p.b), cc = T(p.c);
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = (has_d ? bb * dq[0].v[v] : T(0)) + cc * r.v[v];
                    if (fin_stage) {
                        Cell<T, NV> nh, nl;
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            const float hi = (float)(EPI == EPI_F32CF ? hq[0].v[v] : S.e2.v[v]);
                            const float t = (float)S.e1.v[v] + (float)o.v[v];
                            float sm, er;
                            two_sum(hi, t, sm, er);
                            nh.v[v] = sm;
                            nl.v[v] = er;
                        }
                        if constexpr (NT) store_cell_nt<T, NV>(OUT2, cidx, nl);
                        else store_cell<T, NV>(OUT2, cidx, nl);
                        o = nh;
                    }
                }
                if (p.nf) finite_acc<T, NV>(nfbad, o);
                if constexpr (NT) store_cell_nt<T, NV>(OUT, cidx, o);
                else store_cell<T, NV>(OUT, cidx, o);
            }
            issue(S, j + PF);   // refill this slot for plane j + PF
        }
    }
    watch_flush(p.nf, nfbad);
}

// Centre-plane march of the same tile: the own column's z-queue holds planes
// z-3 .. z+3 before plane z is computed, and plane z (with its x/y halo) is
// staged in LDS and finished in the same iteration.  Fewer live registers
// than the lead-plane march (no queue of in-plane parts), at the price of
// halo loads one plane ahead of the neighbours' interior loads.  The column
// (4 planes ahead) and halo (1 plane ahead) loads keep their raw words until
// the end of the iteration, so their latency overlaps the iteration's work.
// NT: epilogue operands and results stream past L2 (nontemporal), leaving L2
// to the stage input that neighbouring tiles re-read as halo.
//
// NB = 2: the plane is staged in a double buffer and its halo is loaded one
// plane ahead.  NB = 5 (halo ring): the halo of plane z+4 is loaded together
// with the column's plane z+4, i.e. at the iteration in which the
// neighbouring tiles load those cells as their interior, so the shared rows
// are still in L2; it waits in its slot of a 5-plane LDS ring until plane z+4
// is computed.
template <typename T, int SYS, int BX, int BY, int EPI, bool LATE_EPI, bool WALLS, bool NT, int NB>
__device__ __forceinline__ void linear_tile_centre(const LinArgs &p, int x0, int y0, int zb,
                                                   T (&lds)[NB][LinTile<T, SYS, BX, BY>::NS][BY + 6][BX + 6])
{
    static_assert(NB == 2 || NB == 5, "double buffer or 5-plane halo ring");
    constexpr bool HRING = NB == 5;
    using Tile = LinTile<T, SYS, BX, BY>;
    constexpr int NV = Tile::NV, NS = Tile::NS;
    static_assert(BX * BY >= Tile::NHALO, "every halo cell of a plane needs a thread of its own");

    const T *__restrict__ Q = static_cast<const T *>(p.q);
    const T *__restrict__ D = static_cast<const T *>(p.d);
    const T *__restrict__ U0 = static_cast<const T *>(p.u0);
    const T *__restrict__ LO = static_cast<const T *>(p.lo);
    T *__restrict__ OUT = static_cast<T *>(p.out);
    T *__restrict__ OUT2 = static_cast<T *>(p.out2);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int ze = min(zb + p.zc, nz);
    const bool active = (x0 + tx < nx) && (y0 + ty < ny);
    const size_t plane = (size_t)nx * ny;
    const int bcx = WALLS ? p.bc[0] : 0, bcy = WALLS ? p.bc[1] : 0, bcz = WALLS ? p.bc[2] : 0;
    constexpr bool F32C = EPI == EPI_F32C || EPI == EPI_F32CF;
    const bool has_d = F32C && D != nullptr;
    const bool fin_stage = EPI == EPI_F32CF || (EPI == EPI_F32C && p.final_stage);

    float sgx = 1.f, sgy = 1.f;
    const int mx = map_axis(x0 + tx, nx, bcx, sgx);
    const int my = map_axis(y0 + ty, ny, bcy, sgy);
    const size_t col = (size_t)my * nx + mx;

    int hx = -1, hy = -1;
    size_t hcol = 0;
    float hsx = 1.f, hsy = 1.f;
    if (tid < Tile::NHALO) {
        int t = tid;
        if (t < 6 * BY) {
            int row = t / 6, k = t % 6;
            hy = row + 3;
            hx = k < 3 ? k : BX + k;
        } else {
            t -= 6 * BY;
            int colx = t % BX, k = t / BX;
            hx = colx + 3;
            hy = k < 3 ? k : BY + k;
        }
        int gx = map_axis(x0 + hx - 3, nx, bcx, hsx);
        int gy = map_axis(y0 + hy - 3, ny, bcy, hsy);
        hcol = (size_t)gy * nx + gx;
    }
    auto zmap = [&](int z, float &s) -> size_t {
        s = 1.f;
        return (size_t)map_axis(z, nz, bcz, s) * plane;
    };
    struct Raw {
        Cell<T, NV> q, d;
        float s;
    };
    auto issue = [&](Raw &r, size_t c, int z) {
        const size_t off = zmap(z, r.s);
        r.q = load_cell<T, NV>(Q, off + c);
        if (has_d) r.d = load_cell<T, NV>(D, off + c);
    };
    auto apply = [&](const Raw &r, float sx, float sy) -> Cell<T, NV> {
        Cell<T, NV> c = r.q;
        if (has_d) {
#pragma unroll
            for (int v = 0; v < NV; ++v) c.v[v] += r.d.v[v];
        }
        if constexpr (WALLS) apply_bc_sign<T, NV>(c, sx, sy, r.s);
        return c;
    };

    Cell<T, NV> qz[7];
#pragma unroll
    for (int d = -3; d <= 3; ++d) {
        Raw r;
        issue(r, col, zb + d);
        qz[d + 3] = apply(r, sgx, sgy);
    }
    Cell<T, NV> hv;
    if constexpr (HRING) {
        // halos of planes zb .. zb+3 into ring slots 0..3 (visible after the
        // first barrier of the march)
        if (hx >= 0) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                if (zb + d < ze) {
                    Raw r;
                    issue(r, hcol, zb + d);
                    const Cell<T, NV> c = apply(r, hsx, hsy);
#pragma unroll
                    for (int v = 0; v < NS; ++v) lds[d][v][hy][hx] = c.v[v];
                }
            }
        }
    } else if (hx >= 0) {
        Raw r;
        issue(r, hcol, zb);
        hv = apply(r, hsx, hsy);
    }
    // Drain the prologue loads before the march.  Otherwise the queue and
    // halo registers enter the loop as pending load targets, and the
    // compiler's conservative merge at the loop header makes every iteration
    // wait for its own 4-planes-ahead prefetch before the barrier.
    __builtin_amdgcn_s_waitcnt(0);

    const T inv48h = T(p.inv48h), Kc = T(p.K), irho = T(p.invrho), sLLF = T(p.s);
    const T lap = T(p.lapc);

    bool nfbad = false;               // non-finite watch of this thread's outputs
    int buf = 0;                      // slot of plane z
    int hslot = HRING ? 4 : 0;        // ring slot of plane z + 4
    for (int z = zb; z < ze; ++z) {
        // column plane z+4; halo plane z+4 (ring) or z+1
        Raw cq, ch;
        issue(cq, col, z + 4);
        if (HRING ? (hx >= 0 && z + 4 < ze) : (hx >= 0)) issue(ch, hcol, HRING ? z + 4 : z + 1);
        const size_t idx = (size_t)z * plane + col;
        Cell<T, NV> e0, e1, e2;
        auto epi_loads = [&]() {
            if constexpr (EPI == EPI_COMB) e0 = NT ? load_cell_nt<T, NV>(U0, idx) : load_cell<T, NV>(U0, idx);
            if constexpr (F32C) {
                if (has_d) e0 = load_cell<T, NV>(D, idx);
                if (fin_stage) {
                    e1 = load_cell<T, NV>(LO, idx);
                    e2 = load_cell<T, NV>(Q, idx);
                }
            }
        };
        // epilogue operands: issued with the plane loads, or just before use
        // (fewer live registers across the stencil)
        if (!LATE_EPI && active) epi_loads();
#pragma unroll
        for (int v = 0; v < NS; ++v) lds[buf][v][ty + 3][tx + 3] = qz[3].v[v];
        if (!HRING && hx >= 0) {
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
                const T ym3 = lds[buf][v][ty][tx + 3], ym2 = lds[buf][v][ty + 1][tx + 3];
                const T ym1 = lds[buf][v][ty + 2][tx + 3];
                const T yp1 = lds[buf][v][ty + 4][tx + 3], yp2 = lds[buf][v][ty + 5][tx + 3];
                const T yp3 = lds[buf][v][ty + 6][tx + 3];
                const T q0 = qz[3].v[v];
                accB[v] = stencil_B7(xm3, xm2, xm1, q0, xp1, xp2, xp3) + stencil_B7(ym3, ym2, ym1, q0, yp1, yp2, yp3) +
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
            const T L = stencil_L7(row[-3], row[-2], row[-1], q0, row[1], row[2], row[3]) +
                        stencil_L7(lds[buf][0][ty][tx + 3], lds[buf][0][ty + 1][tx + 3], lds[buf][0][ty + 2][tx + 3], q0,
                                   lds[buf][0][ty + 4][tx + 3], lds[buf][0][ty + 5][tx + 3], lds[buf][0][ty + 6][tx + 3]) +
                        stencil_L7(qz[0].v[0], qz[1].v[0], qz[2].v[0], q0, qz[4].v[0], qz[5].v[0], qz[6].v[0]);
            r.v[0] = qz[3].v[1];     // du/dt = v (rhs_function.c:330)
            r.v[1] = lap * L;
        }

        if (active) {
            if (LATE_EPI) epi_loads();
            Cell<T, NV> o;
            if constexpr (EPI == EPI_RHS) {
                o = r;
            } else if constexpr (EPI == EPI_AXPY) {
                const T cc = T(p.c);
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = qz[3].v[v] + cc * r.v[v];
            } else if constexpr (EPI == EPI_COMB) {
                const T aa = T(p.a), bb = T(p.b), cc = T(p.c);
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = aa * e0.v[v] + bb * qz[3].v[v] + cc * r.v[v];
            } else {
                const T bb = T(p.b), cc = T(p.c);
#pragma unroll
                for (int v = 0; v < NV; ++v) o.v[v] = (has_d ? bb * e0.v[v] : T(0)) + cc * r.v[v];
                if (fin_stage) {
                    Cell<T, NV> nh, nl;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const float t = (float)e1.v[v] + (float)o.v[v];
                        float sm, er;
                        two_sum((float)e2.v[v], t, sm, er);
                        nh.v[v] = sm;
                        nl.v[v] = er;
                    }
                    store_cell<T, NV>(OUT2, idx, nl);
                    o = nh;
                }
            }
            if (p.nf) finite_acc<T, NV>(nfbad, o);
            if constexpr (NT) store_cell_nt<T, NV>(OUT, idx, o);
            else store_cell<T, NV>(OUT, idx, o);
        }

#pragma unroll
        for (int k = 0; k < 6; ++k) qz[k] = qz[k + 1];
        qz[6] = apply(cq, sgx, sgy);
        if constexpr (HRING) {
            // slot of plane z+4 = slot of plane z-1, whose readers all passed
            // this iteration's barrier
            if (hx >= 0 && z + 4 < ze) {
                const Cell<T, NV> c = apply(ch, hsx, hsy);
#pragma unroll
                for (int v = 0; v < NS; ++v) lds[hslot][v][hy][hx] = c.v[v];
            }
            buf = buf == NB - 1 ? 0 : buf + 1;
            hslot = hslot == NB - 1 ? 0 : hslot + 1;
        } else {
            if (hx >= 0) hv = apply(ch, hsx, hsy);
            buf ^= 1;
        }
    }
    watch_flush(p.nf, nfbad);
}

// ---------------------------------------------------------------------------
// LDS-DMA march (FP64 ACOUSTIC4, all-periodic grids, RHS and AXPY stages).
//
// The same tile (32 x 16 columns) and arithmetic as the centre-plane march,
// with the plane traffic moved out of registers: every plane of the tile
// with its 3-cell x/y halo (22 x 38 cells, corners included) is copied from
// HBM straight into an LDS ring slot by direct-to-LDS loads (16 bytes per
// lane, global_load_lds_dwordx4), six slots deep (160.5 KB).  At iteration z
// the slots hold planes z .. z+5: plane z is the in-plane stencil's, planes
// z+1 .. z+3 have landed, z+4 and z+5 are in flight.  The z-stencil keeps
// the own column's planes z-3 .. z+3 in registers, the newest read from the
// slot of plane z+3 once it has landed.  So each plane's loads are issued
// five iterations before its stencil and two before its column value is
// needed, and two planes per CU are in flight while a plane is computed,
// without holding them in registers.  (Measured on MI355X at 256^3: loading
// each slot again right after a second barrier in the middle of the
// iteration, one plane deeper, and issuing the stores before the loads were
// both slower.)
//
// The loads are issued from inline assembly so that the compiler does not
// guard every LDS read with a wait for all of them; the waits are explicit:
// loads complete in issue order, and at iteration z the only loads issued
// after plane z+3's are those of plane z+4 (nj per wavefront), so vmcnt <= nj
// means plane z+3 has landed in this wavefront's part (stores count too,
// which can only lengthen the wait); the barrier that follows makes every
// wavefront's part visible and ends all reads of the slot the next loads
// overwrite.  Plane layout in a slot: two
// arrays of 22 x 38 16-byte entries, the (p, u) halves then the (v, w)
// halves of the cells, row-major; one wavefront load fills 64 consecutive
// entries.
struct DmaPlan {
    static constexpr int BX = 32, BY = 16, RW = BX + 6, RH = BY + 6;
    static constexpr int NE = RW * RH;              // 836 cells per plane
    static constexpr int NENT = 2 * NE;             // 1672 sixteen-byte entries
    static constexpr int SLOT = NENT * 16;          // 26752 bytes
    static constexpr int NSLOT = 6;
    static constexpr int NCHUNK = (NENT + 63) / 64; // 27 wavefront loads per plane
    static constexpr int NWAVE = BX * BY / 64;      // 8
    static constexpr int MAXJ = (NCHUNK + NWAVE - 1) / NWAVE;   // 4 loads per wavefront at most
};

template <int EPI, bool RS>
__device__ __forceinline__ void linear_tile_dma(const LinArgs &p, int x0, int y0, int zb, unsigned char *ring)
{
    using D = DmaPlan;
    static_assert(EPI == EPI_RHS || EPI == EPI_AXPY, "LDS-DMA march: RHS and AXPY stages");
    const double *__restrict__ Q = static_cast<const double *>(p.q);
    double *__restrict__ OUT = static_cast<double *>(p.out);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * D::BX + tx;
    const int wave = tid >> 6, lane = tid & 63;
    const int ze = min(zb + p.zc, nz);
    const bool active = (x0 + tx < nx) && (y0 + ty < ny);
    const size_t plane = (size_t)nx * ny;

    // this lane's entries of a plane: chunk q = wave + 8 j covers entries 64 q .. 64 q + 63
    const int nj = wave + D::NWAVE * (D::MAXJ - 1) < D::NCHUNK ? D::MAXJ : D::MAXJ - 1;
    uint32_t eoff[D::MAXJ];
    bool eok[D::MAXJ];
#pragma unroll
    for (int j = 0; j < D::MAXJ; ++j) {
        const int e = (wave + D::NWAVE * j) * 64 + lane;
        eok[j] = j < nj && e < D::NENT;
        const int h = e / D::NE, idx = e - h * D::NE, r = idx / D::RW, c = idx - r * D::RW;
        int gx = x0 + c - 3, gy = y0 + r - 3;
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        eoff[j] = eok[j] ? (uint32_t)(((size_t)gy * nx + gx) * 4 + 2 * h) : 0u;
    }
    auto slot_of = [&](int z) { return (z - zb + 3) % D::NSLOT; };
    auto issue = [&](int z, int slot) {
        int zz = z % nz;
        zz = zz < 0 ? zz + nz : zz;
        const double *src = Q + (size_t)zz * plane * 4;
        unsigned char *dst = ring + slot * D::SLOT;
#pragma unroll
        for (int j = 0; j < D::MAXJ; ++j) {
            if (j < nj) {
                // LDS byte offset of the chunk (wave-uniform) into M0; the
                // lane's 16 bytes land at M0 + 16 * lane
                const unsigned lds = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(
                    (__attribute__((address_space(3))) unsigned char *)(dst + (wave + D::NWAVE * j) * 1024)));
                const double *g = src + eoff[j];
                if (eok[j])
                    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                                 :: "s"(lds), "v"(g) : "memory", "m0");
            }
        }
    };
    // s_waitcnt with only vmcnt constrained (gfx9 encoding: expcnt 7,
    // lgkmcnt 15): all but the loads of the `later` most recent planes
    auto wait_planes = [&](int later) {
        if (later == 0) {
            __builtin_amdgcn_s_waitcnt(0x0F70);
        } else if (later == 1) {
            if (nj == 4) __builtin_amdgcn_s_waitcnt(0x0F74);
            else __builtin_amdgcn_s_waitcnt(0x0F73);
        } else {
            if (nj == 4) __builtin_amdgcn_s_waitcnt(0x0F78);
            else __builtin_amdgcn_s_waitcnt(0x0F76);
        }
    };
    auto cell_at = [&](int slot, int r, int c) -> Cell<double, 4> {
        const unsigned char *b = ring + slot * D::SLOT;
        const double2 a = *reinterpret_cast<const double2 *>(b + (r * D::RW + c) * 16);
        const double2 v = *reinterpret_cast<const double2 *>(b + (D::NE + r * D::RW + c) * 16);
        Cell<double, 4> o;
        o.v[0] = a.x; o.v[1] = a.y; o.v[2] = v.x; o.v[3] = v.y;
        return o;
    };
    auto half_at = [&](int slot, int h, int r, int c) -> double2 {
        return *reinterpret_cast<const double2 *>(ring + slot * D::SLOT + (h * D::NE + r * D::RW + c) * 16);
    };

    // prologue: planes zb-3 .. zb+2 into slots 0 .. 5, their own-column values
    // into the z-queue, then planes zb+3 and zb+4 into the slots of zb-3, zb-2
    for (int d = 0; d < D::NSLOT; ++d) issue(zb - 3 + d, d);
    wait_planes(0);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    Cell<double, 4> qz[7];
#pragma unroll
    for (int d = 0; d < 6; ++d) qz[d + 1] = cell_at(d, ty + 3, tx + 3);
    qz[0] = qz[1];
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue(zb + 3, slot_of(zb + 3));
    if (zb + 4 <= ze + 2) issue(zb + 4, slot_of(zb + 4));

    const double inv48h = p.inv48h, Kc = p.K, irho = p.invrho, sLLF = p.s;
    // Results: RS false, each lane stores its own cell (two 16-byte stores,
    // (p, u) then (v, w)).  RS true, row-complete stores: lanes 0-31 of a
    // wavefront hold row y0 + 2 wave of the tile, lanes 32-63 the next row;
    // after one lane-half exchange (permlane32_swap) the first store writes
    // the first row whole (lane l: cell l & 31, half l >> 5) and the second
    // store the second row, so each 64-byte line of the output is written by
    // one instruction.  Measured on MI355X: the output bytes equal the
    // compulsory ones (7 % more with RS false, some lines reach memory
    // twice); 3.4 % faster at 384^3 (several rounds of workgroups per CU),
    // 0.8 % slower at 256^3 (one round).
    const size_t col = (size_t)(y0 + ty) * nx + (x0 + tx);
    bool nfbad = false;
    for (int z = zb; z < ze; ++z) {
        // plane z+3 landed: the loads issued after its loads are plane z+4's
        // (when issued: planes beyond ze+2 are not needed)
        wait_planes(z + 4 <= ze + 2);
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // every wavefront is past iteration z-1: its slot takes plane z+5
        if (z + 5 <= ze + 2) issue(z + 5, slot_of(z + 5));
        const int s0 = slot_of(z);
#pragma unroll
        for (int k = 0; k < 6; ++k) qz[k] = qz[k + 1];
        qz[6] = cell_at(slot_of(z + 3), ty + 3, tx + 3);

        // in-plane stencils from the slot of plane z, one 16-byte half (two
        // components) at a time; z-stencils from the queue
        double accB[4], dAp[3], dAvel = 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double2 xr[7], yc[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                xr[k] = half_at(s0, h, ty + 3, tx + k);
                yc[k] = half_at(s0, h, ty + k, tx + 3);
            }
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int v = 2 * h + w;
                auto X = [&](int k) { return w ? xr[k].y : xr[k].x; };
                auto Y = [&](int k) { return w ? yc[k].y : yc[k].x; };
                const double q0 = qz[3].v[v];
                accB[v] = stencil_B7(X(0), X(1), X(2), q0, X(4), X(5), X(6)) +
                          stencil_B7(Y(0), Y(1), Y(2), q0, Y(4), Y(5), Y(6)) +
                          stencil_B7(qz[0].v[v], qz[1].v[v], qz[2].v[v], q0, qz[4].v[v], qz[5].v[v], qz[6].v[v]);
                if (v == 0) {
                    dAp[0] = stencil_A7(X(0), X(1), X(2), X(4), X(5), X(6));
                    dAp[1] = stencil_A7(Y(0), Y(1), Y(2), Y(4), Y(5), Y(6));
                    dAp[2] = stencil_A7(qz[0].v[0], qz[1].v[0], qz[2].v[0], qz[4].v[0], qz[5].v[0], qz[6].v[0]);
                } else if (v == 1) {
                    dAvel += stencil_A7(X(0), X(1), X(2), X(4), X(5), X(6));
                } else if (v == 2) {
                    dAvel += stencil_A7(Y(0), Y(1), Y(2), Y(4), Y(5), Y(6));
                } else {
                    dAvel += stencil_A7(qz[0].v[3], qz[1].v[3], qz[2].v[3], qz[4].v[3], qz[5].v[3], qz[6].v[3]);
                }
            }
        }
        Cell<double, 4> r;
        r.v[0] = inv48h * (sLLF * accB[0] - Kc * dAvel);
        r.v[1] = inv48h * (sLLF * accB[1] - irho * dAp[0]);
        r.v[2] = inv48h * (sLLF * accB[2] - irho * dAp[1]);
        r.v[3] = inv48h * (sLLF * accB[3] - irho * dAp[2]);
        if constexpr (RS) {
            Cell<double, 4> o;
            if constexpr (EPI == EPI_RHS) {
                o = r;
            } else {
                const double cc = p.c;
#pragma unroll
                for (int v = 0; v < 4; ++v) o.v[v] = qz[3].v[v] + cc * r.v[v];
            }
            if (active && p.nf) finite_acc<double, 4>(nfbad, o);
            // (p, u) of lanes 32-63 <-> (v, w) of lanes 0-31: all lanes take part
            swap_halves32(o.v[0], o.v[2]);
            swap_halves32(o.v[1], o.v[3]);
            const int xc = x0 + (lane & 31), rowA = y0 + 2 * wave;
            double *outA = OUT + ((size_t)z * plane + (size_t)rowA * nx + xc) * 4 + 2 * (lane >> 5);
            if (xc < nx && rowA < ny)
                __builtin_nontemporal_store(acx_dbl2{o.v[0], o.v[1]}, reinterpret_cast<acx_dbl2 *>(outA));
            if (xc < nx && rowA + 1 < ny)
                __builtin_nontemporal_store(acx_dbl2{o.v[2], o.v[3]}, reinterpret_cast<acx_dbl2 *>(outA + (size_t)nx * 4));
        } else if (active) {
            Cell<double, 4> o;
            if constexpr (EPI == EPI_RHS) {
                o = r;
            } else {
                const double cc = p.c;
#pragma unroll
                for (int v = 0; v < 4; ++v) o.v[v] = qz[3].v[v] + cc * r.v[v];
            }
            if (p.nf) finite_acc<double, 4>(nfbad, o);
            store_cell_nt<double, 4>(OUT, (size_t)z * plane + col, o);
        }
    }
    // every load issued has been waited for by the last iteration (vmcnt 0)
    watch_flush(p.nf, nfbad);
}

// The LDS-DMA march for the FP64 WAVE2 system (u, v: 16 bytes per cell, one
// slot entry), RHS and AXPY stages on all-periodic grids: 6 slots of 13.4 KB,
// so two 512-thread blocks fit a CU.  du/dt = v, dv/dt = c^2 Lap(u) with the
// L7 stencil per axis (3d/weno4/plane_wave/rhs_function.c:326-338 with the
// shipped linear weights).
struct DmaPlanW2 {
    static constexpr int NE = DmaPlan::NE;                       // 836 cells per plane
    static constexpr int SLOT = NE * 16;                         // 13376 bytes
    static constexpr int NSLOT = 6;
    static constexpr int NCHUNK = (NE + 63) / 64;                // 14 wavefront loads per plane
    static constexpr int MAXJ = (NCHUNK + DmaPlan::NWAVE - 1) / DmaPlan::NWAVE;   // 2
};

template <int EPI>
__device__ __forceinline__ void linear_tile_dma_w2(const LinArgs &p, int x0, int y0, int zb, unsigned char *ring)
{
    using D = DmaPlan;
    using W = DmaPlanW2;
    static_assert(EPI == EPI_RHS || EPI == EPI_AXPY, "LDS-DMA march: RHS and AXPY stages");
    const double *__restrict__ Q = static_cast<const double *>(p.q);
    double *__restrict__ OUT = static_cast<double *>(p.out);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * D::BX + tx;
    const int wave = tid >> 6, lane = tid & 63;
    const int ze = min(zb + p.zc, nz);
    const bool active = (x0 + tx < nx) && (y0 + ty < ny);
    const size_t plane = (size_t)nx * ny;

    const int nj = wave + D::NWAVE * (W::MAXJ - 1) < W::NCHUNK ? W::MAXJ : W::MAXJ - 1;
    uint32_t eoff[W::MAXJ];
    bool eok[W::MAXJ];
#pragma unroll
    for (int j = 0; j < W::MAXJ; ++j) {
        const int e = (wave + D::NWAVE * j) * 64 + lane;
        eok[j] = j < nj && e < W::NE;
        const int r = e / D::RW, c = e - r * D::RW;
        int gx = x0 + c - 3, gy = y0 + r - 3;
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        eoff[j] = eok[j] ? (uint32_t)(((size_t)gy * nx + gx) * 2) : 0u;
    }
    auto slot_of = [&](int z) { return (z - zb + 3) % W::NSLOT; };
    auto issue = [&](int z, int slot) {
        int zz = z % nz;
        zz = zz < 0 ? zz + nz : zz;
        const double *src = Q + (size_t)zz * plane * 2;
        unsigned char *dst = ring + slot * W::SLOT;
#pragma unroll
        for (int j = 0; j < W::MAXJ; ++j) {
            if (j < nj) {
                const unsigned lds = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(
                    (__attribute__((address_space(3))) unsigned char *)(dst + (wave + D::NWAVE * j) * 1024)));
                const double *g = src + eoff[j];
                if (eok[j])
                    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                                 :: "s"(lds), "v"(g) : "memory", "m0");
            }
        }
    };
    // vmcnt: all loads, or all but the nj of the most recent plane
    auto wait_planes = [&](bool keep_last) {
        if (!keep_last) __builtin_amdgcn_s_waitcnt(0x0F70);
        else if (nj == 2) __builtin_amdgcn_s_waitcnt(0x0F72);
        else __builtin_amdgcn_s_waitcnt(0x0F71);
    };
    auto cell_at = [&](int slot, int r, int c) -> double2 {
        return *reinterpret_cast<const double2 *>(ring + slot * W::SLOT + (r * D::RW + c) * 16);
    };
    auto u_at = [&](int slot, int r, int c) -> double {
        return *reinterpret_cast<const double *>(ring + slot * W::SLOT + (r * D::RW + c) * 16);
    };

    for (int d = 0; d < W::NSLOT; ++d) issue(zb - 3 + d, d);
    wait_planes(false);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    double2 qz[7];
#pragma unroll
    for (int d = 0; d < 6; ++d) qz[d + 1] = cell_at(d, ty + 3, tx + 3);
    qz[0] = qz[1];
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue(zb + 3, slot_of(zb + 3));
    if (zb + 4 <= ze + 2) issue(zb + 4, slot_of(zb + 4));

    const double lap = p.lapc;
    const size_t col = (size_t)(y0 + ty) * nx + (x0 + tx);
    bool nfbad = false;
    for (int z = zb; z < ze; ++z) {
        wait_planes(z + 4 <= ze + 2);
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (z + 5 <= ze + 2) issue(z + 5, slot_of(z + 5));
        const int s0 = slot_of(z);
#pragma unroll
        for (int k = 0; k < 6; ++k) qz[k] = qz[k + 1];
        qz[6] = cell_at(slot_of(z + 3), ty + 3, tx + 3);
        const double q0 = qz[3].x;
        const double lx = stencil_L7(u_at(s0, ty + 3, tx), u_at(s0, ty + 3, tx + 1), u_at(s0, ty + 3, tx + 2), q0,
                                     u_at(s0, ty + 3, tx + 4), u_at(s0, ty + 3, tx + 5), u_at(s0, ty + 3, tx + 6));
        const double ly = stencil_L7(u_at(s0, ty, tx + 3), u_at(s0, ty + 1, tx + 3), u_at(s0, ty + 2, tx + 3), q0,
                                     u_at(s0, ty + 4, tx + 3), u_at(s0, ty + 5, tx + 3), u_at(s0, ty + 6, tx + 3));
        const double lz = stencil_L7(qz[0].x, qz[1].x, qz[2].x, q0, qz[4].x, qz[5].x, qz[6].x);
        Cell<double, 2> r;
        r.v[0] = qz[3].y;                 // du/dt = v
        r.v[1] = lap * ((lx + ly) + lz);
        if (active) {
            Cell<double, 2> o;
            if constexpr (EPI == EPI_RHS) {
                o = r;
            } else {
                const double cc = p.c;
                o.v[0] = qz[3].x + cc * r.v[0];
                o.v[1] = qz[3].y + cc * r.v[1];
            }
            if (p.nf) finite_acc<double, 2>(nfbad, o);
            __builtin_nontemporal_store(acx_dbl2{o.v[0], o.v[1]},
                                        reinterpret_cast<acx_dbl2 *>(OUT + ((size_t)z * plane + col) * 2));
        }
    }
    watch_flush(p.nf, nfbad);
}

} // namespace acx
