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
// chunk of z-planes.  x-neighbour face states and fluxes move by lane
// shuffles inside a 32-lane row, y-neighbours through LDS, z-neighbours stay
// in registers.  Selective mode (per-cell flags) replaces the full CWENO by
// the linear 4th-order polynomial wherever the flag bit of a component is 0.
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

enum { EPI_RHS = 0, EPI_AXPY = 1, EPI_COMB = 2, EPI_F32C = 3 };

template <typename T> struct Axis3 { T c1[3], c2[3], c3[3]; };

template <typename T>
__device__ __forceinline__ T pos_pow4_inv(T g, T t)
{
    T t2 = t * t;
    return g / (t2 * t2);
}

// Linear 4th-order polynomial along each axis (reconstruction.c:90-101)
template <typename T>
__device__ __forceinline__ void axis_linear(T m2, T m1, T u0, T p1, T p2, T &c1, T &c2, T &c3)
{
    c1 = T(1.0 / 120.0) * (T(11) * m2 - T(82) * m1 + T(82) * p1 - T(11) * p2);
    c2 = T(0.5) * (m1 - T(2) * u0 + p1);
    c3 = T(1.0 / 12.0) * (-m2 + T(2) * (m1 - p1) + p2);
}

// Full CWENO4: returns the axis coefficients (c1,c2,c3 per axis) and wt_ratio.
template <typename T>
__device__ __forceinline__ T cweno_axes(const T (&s)[33], const CwArgs &p, Axis3<T> &o)
{
    // stencil slots (see gather order in the kernel)
    const T u0 = s[0];
    const T xm2 = s[1], xm1 = s[2], xp1 = s[3], xp2 = s[4];
    const T ym2 = s[5], ym1 = s[6], yp1 = s[7], yp2 = s[8];
    const T zm2 = s[9], zm1 = s[10], zp1 = s[11], zp2 = s[12];
    const T xy_pp = s[13], xy_pm = s[14], xy_mp = s[15], xy_mm = s[16];
    const T yz_pp = s[17], yz_pm = s[18], yz_mp = s[19], yz_mm = s[20];
    const T zx_pp = s[21], zx_pm = s[22], zx_mp = s[23], zx_mm = s[24];
    const T ppp = s[25], mpp = s[26], pmp = s[27], ppm = s[28];
    const T mmp = s[29], mpm = s[30], pmm = s[31], mmm = s[32];

    // R4 (reconstruction.c:90-108)
    T ux, uy, uz, uxx, uyy, uzz, uxxx, uyyy, uzzz;
    axis_linear(xm2, xm1, u0, xp1, xp2, ux, uxx, uxxx);
    axis_linear(ym2, ym1, u0, yp1, yp2, uy, uyy, uyyy);
    axis_linear(zm2, zm1, u0, zp1, zp2, uz, uzz, uzzz);
    const T uxy = T(0.25) * (xy_mm - xy_mp - xy_pm + xy_pp);
    const T uyz = T(0.25) * (yz_mm - yz_mp - yz_pm + yz_pp);
    const T uxz = T(0.25) * (zx_mm - zx_mp - zx_pm + zx_pp);
    const T uxxy = T(0.25) * (-xy_mm + xy_mp - xy_pm + xy_pp) + T(0.5) * (ym1 - yp1);
    const T uxyy = T(0.5) * (xm1 - xp1) + T(0.25) * (xy_pp - xy_mm - xy_mp + xy_pm);
    const T uyyz = T(0.25) * (-yz_mm + yz_mp - yz_pm + yz_pp) + T(0.5) * (zm1 - zp1);
    const T uyzz = T(0.5) * (ym1 - yp1) + T(0.25) * (yz_pp - yz_mm - yz_mp + yz_pm);
    const T uxxz = T(0.25) * (-zx_mm + zx_pm - zx_mp + zx_pp) + T(0.5) * (zm1 - zp1);
    const T uxzz = T(0.5) * (xm1 - xp1) + T(0.25) * (zx_pp - zx_mm - zx_pm + zx_mp);
    const T uxyz = T(0.125) * (ppp - mpp - pmp - ppm + mmp + mpm + pmm - mmm);

    const T k13 = T(13.0 / 3.0), k7 = T(7.0 / 6.0);
    const T is4 = ux * ux + uy * uy + uz * uz + k13 * (uxx * uxx + uyy * uyy + uzz * uzz) +
                  k7 * (uxy * uxy + uyz * uyz + uxz * uxz) + T(0.2) * (ux * uxxx + uy * uyyy + uz * uzzz) +
                  T(4.7) * (uxxy * uxxy + uxxz * uxxz + uxyy * uxyy + uxzz * uxzz + uyyz * uyyz + uyzz * uyzz) +
                  T(39.06) * (uxxx * uxxx + uyyy * uyyy + uzzz * uzzz) + T(61.0 / 48.0) * uxyz * uxyz;

    // R3 stencils (reconstruction.c:121-203): central, +x, -x, +y, -y, +z, -z.
    // Central derivatives coincide with the R4 ones for xx, yy, zz, xy, yz, xz.
    const T cx = T(0.5) * (xp1 - xm1), cy = T(0.5) * (yp1 - ym1), cz = T(0.5) * (zp1 - zm1);
    const T dx1 = T(-1.5) * u0 + T(2) * xp1 - T(0.5) * xp2, dxx1 = T(0.5) * (u0 - T(2) * xp1 + xp2);
    const T dxy1 = T(0.5) * (xy_pp - xy_pm) - cy, dxz1 = T(0.5) * (zx_pp - zx_mp) - cz;
    const T dx2 = T(-2) * xm1 + T(0.5) * xm2 + T(1.5) * u0, dxx2 = T(0.5) * (xm2 - T(2) * xm1 + u0);
    const T dxy2 = T(0.5) * (xy_mm - xy_mp) + cy, dxz2 = T(0.5) * (zx_mm - zx_pm) + cz;
    const T dy3 = T(-1.5) * u0 + T(2) * yp1 - T(0.5) * yp2, dyy3 = T(0.5) * (u0 - T(2) * yp1 + yp2);
    const T dxy3 = T(0.5) * (xy_pp - xy_mp) - cx, dyz3 = T(0.5) * (yz_pp - yz_pm) - cz;
    const T dy4 = T(-2) * ym1 + T(0.5) * ym2 + T(1.5) * u0, dyy4 = T(0.5) * (ym2 - T(2) * ym1 + u0);
    const T dxy4 = T(0.5) * (xy_mm - xy_pm) + cx, dyz4 = T(0.5) * (yz_mm - yz_mp) + cz;
    const T dz5 = T(-1.5) * u0 + T(2) * zp1 - T(0.5) * zp2, dzz5 = T(0.5) * (u0 - T(2) * zp1 + zp2);
    const T dyz5 = T(0.5) * (yz_pp - yz_mp) - cy, dxz5 = T(0.5) * (zx_pp - zx_pm) - cx;
    const T dz6 = T(-2) * zm1 + T(0.5) * zm2 + T(1.5) * u0, dzz6 = T(0.5) * (zm2 - T(2) * zm1 + u0);
    const T dyz6 = T(0.5) * (yz_mm - yz_pm) + cy, dxz6 = T(0.5) * (zx_mm - zx_mp) + cx;

    const T Cx = cx * cx + k13 * uxx * uxx, Cy = cy * cy + k13 * uyy * uyy, Cz = cz * cz + k13 * uzz * uzz;
    const T Cxy = k7 * uxy * uxy, Cyz = k7 * uyz * uyz, Cxz = k7 * uxz * uxz;
    T is3[7];
    is3[0] = Cx + Cy + Cz + Cxy + Cyz + Cxz;
    is3[1] = (dx1 * dx1 + k13 * dxx1 * dxx1) + Cy + Cz + k7 * dxy1 * dxy1 + Cyz + k7 * dxz1 * dxz1;
    is3[2] = (dx2 * dx2 + k13 * dxx2 * dxx2) + Cy + Cz + k7 * dxy2 * dxy2 + Cyz + k7 * dxz2 * dxz2;
    is3[3] = Cx + (dy3 * dy3 + k13 * dyy3 * dyy3) + Cz + k7 * dxy3 * dxy3 + k7 * dyz3 * dyz3 + Cxz;
    is3[4] = Cx + (dy4 * dy4 + k13 * dyy4 * dyy4) + Cz + k7 * dxy4 * dxy4 + k7 * dyz4 * dyz4 + Cxz;
    is3[5] = Cx + Cy + (dz5 * dz5 + k13 * dzz5 * dzz5) + Cxy + k7 * dyz5 * dyz5 + k7 * dxz5 * dxz5;
    is3[6] = Cx + Cy + (dz6 * dz6 + k13 * dzz6 * dzz6) + Cxy + k7 * dyz6 * dyz6 + k7 * dxz6 * dxz6;

    // nonlinear weights, p = 4, eps = 1e-12 (reconstruction.c:118,205-218)
    const T eps = T(1.0e-12);
    const T g4 = T(p.g4), g3c = T(p.g3c), g3s = T(p.g3s);
    T w4, w3[7], wt;
    if constexpr (sizeof(T) == 8) {
        w4 = pos_pow4_inv(g4, is4 + eps);
        T total = w4;
        w3[0] = pos_pow4_inv(g3c, is3[0] + eps);
        total += w3[0];
#pragma unroll
        for (int k = 1; k < 7; ++k) {
            w3[k] = pos_pow4_inv(g3s, is3[k] + eps);
            total += w3[k];
        }
        wt = (w4 / total) * T(p.inv_g4);
        const T it = T(1) / total;
#pragma unroll
        for (int k = 0; k < 7; ++k) w3[k] *= it;
    } else {
        // FP32: ratio form gamma_k ((IS_min+eps)/(IS_k+eps))^4, identical after
        // normalisation but free of the underflow of (IS+eps)^4 ~ 1e-48.
        T t4 = is4 + eps, tm = t4, t3[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) { t3[k] = is3[k] + eps; tm = fminf(tm, t3[k]); }
        auto rw = [&](T g, T t) { T r = tm * __builtin_amdgcn_rcpf(t); r = r * r; return g * r * r; };
        w4 = rw(g4, t4);
        T total = w4;
        w3[0] = rw(g3c, t3[0]);
        total += w3[0];
#pragma unroll
        for (int k = 1; k < 7; ++k) { w3[k] = rw(g3s, t3[k]); total += w3[k]; }
        const T it = __builtin_amdgcn_rcpf(total);
        wt = w4 * it * T(p.inv_g4);
#pragma unroll
        for (int k = 0; k < 7; ++k) w3[k] *= it;
    }

    // CWENO combination, reconstruction.c:223-257, for the face-relevant
    // coefficients only.  lin = sum gamma_k R3_k, nl = sum w_k R3_k.
    const T gc = g3c + T(4) * g3s;                // weight of stencils central along an axis
    const T Wx = w3[0] + w3[3] + w3[4] + w3[5] + w3[6];
    const T Wy = w3[0] + w3[1] + w3[2] + w3[5] + w3[6];
    const T Wz = w3[0] + w3[1] + w3[2] + w3[3] + w3[4];
    o.c1[0] = wt * (ux - (gc * cx + g3s * (dx1 + dx2))) + (Wx * cx + w3[1] * dx1 + w3[2] * dx2);
    o.c2[0] = wt * (uxx - (gc * uxx + g3s * (dxx1 + dxx2))) + (Wx * uxx + w3[1] * dxx1 + w3[2] * dxx2);
    o.c3[0] = wt * uxxx;
    o.c1[1] = wt * (uy - (gc * cy + g3s * (dy3 + dy4))) + (Wy * cy + w3[3] * dy3 + w3[4] * dy4);
    o.c2[1] = wt * (uyy - (gc * uyy + g3s * (dyy3 + dyy4))) + (Wy * uyy + w3[3] * dyy3 + w3[4] * dyy4);
    o.c3[1] = wt * uyyy;
    o.c1[2] = wt * (uz - (gc * cz + g3s * (dz5 + dz6))) + (Wz * cz + w3[5] * dz5 + w3[6] * dz6);
    o.c2[2] = wt * (uzz - (gc * uzz + g3s * (dzz5 + dzz6))) + (Wz * uzz + w3[5] * dzz5 + w3[6] * dzz6);
    o.c3[2] = wt * uzzz;
    return wt;
}

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

template <typename T, int SYS, int BX, int BY, int EPI, bool SEL>
__global__ void __launch_bounds__(BX *BY)
k_cweno(CwArgs p)
{
    constexpr int NV = SYS == ACX_ACOUSTIC4 ? 4 : 2;
    constexpr int NS = SYS == ACX_ACOUSTIC4 ? 4 : 1;
    static_assert(BX == 32, "x exchange uses 32-lane row shuffles");
    __shared__ T s_fym[NS][BY][BX];   // -y face state of each column
    __shared__ T s_Fy[NS][BY][BX];    // flux through the +y face of each column

    const T *__restrict__ Q = static_cast<const T *>(p.q);
    const T *__restrict__ D = static_cast<const T *>(p.d);
    const T *__restrict__ U0 = static_cast<const T *>(p.u0);
    const T *__restrict__ LO = static_cast<const T *>(p.lo);
    T *__restrict__ OUT = static_cast<T *>(p.out);
    T *__restrict__ OUT2 = static_cast<T *>(p.out2);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tile = xcd_logical_tile(blockIdx.x, gridDim.x);
    const int gx = (tile % p.ntx) * (BX - 2) - 1 + tx;
    const int gy = ((tile / p.ntx) % p.nty) * (BY - 2) - 1 + ty;
    const int zb = (tile / (p.ntx * p.nty)) * p.zc, ze = min(zb + p.zc, nz);
    const bool interior = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2 && gx < nx && gy < ny;
    const size_t plane = (size_t)nx * ny;
    const int perx = p.bc[0] == ACX_PERIODIC, pery = p.bc[1] == ACX_PERIODIC, perz = p.bc[2] == ACX_PERIODIC;

    // in-plane offsets of the 13 columns the stencil touches, with their ghost signs
    // order: (0,0) (-2,0) (-1,0) (1,0) (2,0) (0,-2) (0,-1) (0,1) (0,2) (1,1) (1,-1) (-1,1) (-1,-1)
    const int cdx[13] = {0, -2, -1, 1, 2, 0, 0, 0, 0, 1, 1, -1, -1};
    const int cdy[13] = {0, 0, 0, 0, 0, -2, -1, 1, 2, 1, -1, 1, -1};
    uint32_t coff[13];
    float csg[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        float s1 = 1.f, s2 = 1.f;
        int mx = map_axis(gx + cdx[k], nx, perx, s1);
        int my = map_axis(gy + cdy[k], ny, pery, s2);
        coff[k] = (uint32_t)(my * nx + mx);
        csg[k] = s1 * s2;
    }

    const T K = T(p.K), irho = T(p.invrho), sL = T(p.s);

    // z-march state
    T zp_prev[NS], Fz_prev[NS], dxy_prev[NS];
#pragma unroll
    for (int v = 0; v < NS; ++v) { zp_prev[v] = T(0); Fz_prev[v] = T(0); dxy_prev[v] = T(0); }
    T vcen_prev = T(0);   // WAVE2: v at the previous plane (du/dt = v)

    for (int m = zb - 1; m <= ze; ++m) {
        // plane offsets m-2..m+2 with ghost signs
        size_t poff[5];
        float psg[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            float s = 1.f;
            poff[k] = (size_t)map_axis(m - 2 + k, nz, perz, s) * plane;
            psg[k] = s;
        }
        uint8_t fl = 0xff;
        if constexpr (SEL) {
            fl = p.flags[poff[2] + coff[0]];   // flag of the physical (mapped) cell
        }

        T fm[NS][3], fp[NS][3];
        T vcen = T(0);
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            auto ld = [&](int pk, int ck) -> T {
                size_t idx = (poff[pk] + coff[ck]) * NV + v;
                T val = Q[idx];
                if constexpr (EPI == EPI_F32C) { if (D) val += D[idx]; }
                if (v == 0) val *= T(psg[pk] * csg[ck]);
                return val;
            };
            T st[33];
            // centre plane m (slot 2): axis arms and xy diagonals
            st[0] = ld(2, 0);
            st[1] = ld(2, 1); st[2] = ld(2, 2); st[3] = ld(2, 3); st[4] = ld(2, 4);
            st[5] = ld(2, 5); st[6] = ld(2, 6); st[7] = ld(2, 7); st[8] = ld(2, 8);
            st[9] = ld(0, 0); st[10] = ld(1, 0); st[11] = ld(3, 0); st[12] = ld(4, 0);
            st[13] = ld(2, 9); st[14] = ld(2, 10); st[15] = ld(2, 11); st[16] = ld(2, 12);
            // yz diagonals: (y,z) = (+,+) (+,-) (-,+) (-,-)
            st[17] = ld(3, 7); st[18] = ld(1, 7); st[19] = ld(3, 6); st[20] = ld(1, 6);
            // zx diagonals named (z,x): (+,+) (+,-) (-,+) (-,-)
            st[21] = ld(3, 3); st[22] = ld(3, 2); st[23] = ld(1, 3); st[24] = ld(1, 2);
            // corners (x,y,z): +++ -++ +-+ ++- --+ -+- +-- ---
            st[25] = ld(3, 9); st[26] = ld(3, 11); st[27] = ld(3, 10); st[28] = ld(1, 9);
            st[29] = ld(3, 12); st[30] = ld(1, 11); st[31] = ld(1, 10); st[32] = ld(1, 12);

            Axis3<T> ax;
            bool full = true;
            if constexpr (SEL) full = (fl >> v) & 1;
            if (full) {
                cweno_axes<T>(st, p, ax);
            } else {
                axis_linear(st[1], st[2], st[0], st[3], st[4], ax.c1[0], ax.c2[0], ax.c3[0]);
                axis_linear(st[5], st[6], st[0], st[7], st[8], ax.c1[1], ax.c2[1], ax.c3[1]);
                axis_linear(st[9], st[10], st[0], st[11], st[12], ax.c1[2], ax.c2[2], ax.c3[2]);
            }
            T m3[3], p3[3];
            face_q<T, SYS>(st[0], ax, m3, p3);
#pragma unroll
            for (int d = 0; d < 3; ++d) { fm[v][d] = m3[d]; fp[v][d] = p3[d]; }
        }
        if constexpr (SYS == ACX_WAVE2) {
            size_t idx = (poff[2] + coff[0]) * NV + 1;
            vcen = Q[idx];
            if constexpr (EPI == EPI_F32C) { if (D) vcen += D[idx]; }
        }

        // ---- y exchange of -y face states through LDS
#pragma unroll
        for (int v = 0; v < NS; ++v) s_fym[v][ty][tx] = fm[v][1];
        __syncthreads();

        // ---- fluxes through the +x face (right neighbour by lane shuffle) and +y face
        T Fx[NS], Fy[NS];
        {
            T L[NS], R[NS];
#pragma unroll
            for (int v = 0; v < NS; ++v) { L[v] = fp[v][0]; R[v] = __shfl_down(fm[v][0], 1, 32); }
            face_flux<T, SYS, NS>(0, L, R, K, irho, sL, Fx);
#pragma unroll
            for (int v = 0; v < NS; ++v) { L[v] = fp[v][1]; R[v] = ty + 1 < BY ? s_fym[v][ty + 1][tx] : T(0); }
            face_flux<T, SYS, NS>(1, L, R, K, irho, sL, Fy);
        }
#pragma unroll
        for (int v = 0; v < NS; ++v) s_Fy[v][ty][tx] = Fy[v];
        __syncthreads();

        T dxy[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            const T Fxm = __shfl_up(Fx[v], 1, 32);
            const T Fym = ty > 0 ? s_Fy[v][ty - 1][tx] : T(0);
            dxy[v] = (Fx[v] - Fxm) + (Fy[v] - Fym);
        }

        // ---- z flux through face m-1/2 and output of plane m-1
        if (m >= zb) {
            T Fz[NS], L[NS], R[NS];
#pragma unroll
            for (int v = 0; v < NS; ++v) { L[v] = zp_prev[v]; R[v] = fm[v][2]; }
            face_flux<T, SYS, NS>(2, L, R, K, irho, sL, Fz);
            if (m >= zb + 1 && interior) {
                const size_t cell = (size_t)(m - 1) * plane + coff[0];
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
                } else if constexpr (EPI == EPI_AXPY) {
                    Cell<T, NV> q = load_cell<T, NV>(Q, cell);
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = q.v[v] + T(p.c) * r.v[v];
                } else if constexpr (EPI == EPI_COMB) {
                    Cell<T, NV> q = load_cell<T, NV>(Q, cell), u0 = load_cell<T, NV>(U0, cell);
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = T(p.a) * u0.v[v] + T(p.b) * q.v[v] + T(p.c) * r.v[v];
                } else {
                    Cell<T, NV> dcur;
                    if (D) dcur = load_cell<T, NV>(D, cell);
                    else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) dcur.v[v] = T(0);
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) o.v[v] = T(p.b) * dcur.v[v] + T(p.c) * r.v[v];
                    if (p.final_stage) {
                        Cell<T, NV> hi = load_cell<T, NV>(Q, cell), lo = load_cell<T, NV>(LO, cell), nh, nl;
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            float t = (float)lo.v[v] + (float)o.v[v];
                            float s, e;
                            two_sum((float)hi.v[v], t, s, e);
                            nh.v[v] = s;
                            nl.v[v] = e;
                        }
                        store_cell<T, NV>(OUT2, cell, nl);
                        o = nh;
                    }
                }
                store_cell<T, NV>(OUT, cell, o);
            }
#pragma unroll
            for (int v = 0; v < NS; ++v) Fz_prev[v] = Fz[v];
        }
#pragma unroll
        for (int v = 0; v < NS; ++v) { zp_prev[v] = fp[v][2]; dxy_prev[v] = dxy[v]; }
        vcen_prev = vcen;
    }
}

// ---------------------------------------------------------------------------

template <typename T, int SYS, int EPI, bool SEL>
static int launch_cw_t(const CwArgs &a, hipStream_t s)
{
    constexpr int BX = 32, BY = 16;
    CwArgs b = a;
    b.ntx = (a.nx + BX - 3) / (BX - 2);
    b.nty = (a.ny + BY - 3) / (BY - 2);
    b.ntz = (a.nz + a.zc - 1) / a.zc;
    dim3 blk(BX, BY, 1);
    dim3 grd(b.ntx * b.nty * b.ntz, 1, 1);
    hipLaunchKernelGGL((k_cweno<T, SYS, BX, BY, EPI, SEL>), grd, blk, 0, s, b);
    return hip_check(hipGetLastError(), "k_cweno launch");
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
