// acx_cweno_core.h -- the CWENO4 reconstruction of one cell and component,
// restricted to what face averages need (shared by the stage kernel, the
// label kernel of the learned indicator and the ensemble kernel).
//
// 3d/weno4/plane_wave/reconstruction.c:38-259: R4 derivatives (:90-108),
// IS_R4 (:110-116), seven R3 stencils (:121-203), weights with p = 4 and
// eps = 1e-12 (:118,205-218), CWENO combination (:223-257).
#pragma once
#include "acx_device.h"
#include "acx_internal.h"

namespace acx {

template <typename T> struct Axis3 { T c1[3], c2[3], c3[3]; };

// 1/t in FP64: hardware reciprocal estimate refined by one Newton step.
// Measured on gfx950 over 2^20 seeded t in [1e-50, 2e50]: the estimate alone
// has |r t - 1| up to 4.6e-8 (2^-24.4), one step brings it to 2.2e-15
// (2^-48.7), which puts the nonlinear weights within about 20 units in the
// last place: far inside the 1e-12 agreement with the FP64 reference
// scheme.  A second step (full precision) cost 2 % of the FP64 nonlinear
// step.
__device__ __forceinline__ double rcp64(double t)
{
    double r = __builtin_amdgcn_rcp(t);
    const double e = fma(-t, r, 1.0);
    return fma(r, e, r);
}

template <typename T>
__device__ __forceinline__ T pos_pow4_inv(T g, T t)
{
    T t2 = t * t;
    return g * rcp64(t2 * t2);
}

// Linear 4th-order polynomial along each axis (reconstruction.c:90-101)
template <typename T>
__device__ __forceinline__ void axis_linear(T m2, T m1, T u0, T p1, T p2, T &c1, T &c2, T &c3)
{
    c1 = T(1.0 / 120.0) * (T(11) * m2 - T(82) * m1 + T(82) * p1 - T(11) * p2);
    c2 = T(0.5) * (m1 - T(2) * u0 + p1);
    c3 = T(1.0 / 12.0) * (-m2 + T(2) * (m1 - p1) + p2);
}

// Smoothness indicators and normalised nonlinear weights of one cell and
// component; get(q) returns stencil slot q (0: centre; 1-4, 5-8, 9-12: the
// x, y, z arms m2 m1 p1 p2; 13-16 xy, 17-20 yz, 21-24 zx diagonals; 25-32
// corners).  wt = wt_ratio (reconstruction.c:236), w3[7] the R3 weights.
// The indicators are accumulated in three groups (axis arms, edge diagonals,
// corners); with LDS-backed values the memory clobbers between the groups keep
// only one group live at a time.
template <typename T, typename Get>
__device__ __forceinline__ void cweno_weights(Get get, const CwArgs &p, T &wt, T (&w3)[7])
{
    const T k13 = T(13.0 / 3.0), k7 = T(7.0 / 6.0);
    T is4, is3[7], cx, cy, cz;
    {   // axis arms: R4 x/y/z derivatives and the sided R3 stencils (reconstruction.c:90-101,121-203)
        const T u0 = get(0);
        const T xm2 = get(1), xm1 = get(2), xp1 = get(3), xp2 = get(4);
        const T ym2 = get(5), ym1 = get(6), yp1 = get(7), yp2 = get(8);
        const T zm2 = get(9), zm1 = get(10), zp1 = get(11), zp2 = get(12);
        T ux, uy, uz, uxx, uyy, uzz, uxxx, uyyy, uzzz;
        axis_linear(xm2, xm1, u0, xp1, xp2, ux, uxx, uxxx);
        axis_linear(ym2, ym1, u0, yp1, yp2, uy, uyy, uyyy);
        axis_linear(zm2, zm1, u0, zp1, zp2, uz, uzz, uzzz);
        is4 = ux * ux + uy * uy + uz * uz + k13 * (uxx * uxx + uyy * uyy + uzz * uzz) +
              T(0.2) * (ux * uxxx + uy * uyyy + uz * uzzz) + T(39.06) * (uxxx * uxxx + uyyy * uyyy + uzzz * uzzz);
        cx = T(0.5) * (xp1 - xm1);
        cy = T(0.5) * (yp1 - ym1);
        cz = T(0.5) * (zp1 - zm1);
        const T dx1 = T(-1.5) * u0 + T(2) * xp1 - T(0.5) * xp2, dxx1 = T(0.5) * (u0 - T(2) * xp1 + xp2);
        const T dx2 = T(-2) * xm1 + T(0.5) * xm2 + T(1.5) * u0, dxx2 = T(0.5) * (xm2 - T(2) * xm1 + u0);
        const T dy3 = T(-1.5) * u0 + T(2) * yp1 - T(0.5) * yp2, dyy3 = T(0.5) * (u0 - T(2) * yp1 + yp2);
        const T dy4 = T(-2) * ym1 + T(0.5) * ym2 + T(1.5) * u0, dyy4 = T(0.5) * (ym2 - T(2) * ym1 + u0);
        const T dz5 = T(-1.5) * u0 + T(2) * zp1 - T(0.5) * zp2, dzz5 = T(0.5) * (u0 - T(2) * zp1 + zp2);
        const T dz6 = T(-2) * zm1 + T(0.5) * zm2 + T(1.5) * u0, dzz6 = T(0.5) * (zm2 - T(2) * zm1 + u0);
        // central R3 shares the R4 second derivatives
        const T Cx = cx * cx + k13 * uxx * uxx, Cy = cy * cy + k13 * uyy * uyy, Cz = cz * cz + k13 * uzz * uzz;
        is3[0] = Cx + Cy + Cz;
        is3[1] = (dx1 * dx1 + k13 * dxx1 * dxx1) + Cy + Cz;
        is3[2] = (dx2 * dx2 + k13 * dxx2 * dxx2) + Cy + Cz;
        is3[3] = Cx + (dy3 * dy3 + k13 * dyy3 * dyy3) + Cz;
        is3[4] = Cx + (dy4 * dy4 + k13 * dyy4 * dyy4) + Cz;
        is3[5] = Cx + Cy + (dz5 * dz5 + k13 * dzz5 * dzz5);
        is3[6] = Cx + Cy + (dz6 * dz6 + k13 * dzz6 * dzz6);
    }
    asm volatile("" ::: "memory");
    {   // edge diagonals: mixed derivatives (reconstruction.c:102-107, sided R3 cross terms)
        const T ym1 = get(6), yp1 = get(7), zm1 = get(10), zp1 = get(11), xm1 = get(2), xp1 = get(3);
        const T xy_pp = get(13), xy_pm = get(14), xy_mp = get(15), xy_mm = get(16);
        const T yz_pp = get(17), yz_pm = get(18), yz_mp = get(19), yz_mm = get(20);
        const T zx_pp = get(21), zx_pm = get(22), zx_mp = get(23), zx_mm = get(24);
        const T uxy = T(0.25) * (xy_mm - xy_mp - xy_pm + xy_pp);
        const T uyz = T(0.25) * (yz_mm - yz_mp - yz_pm + yz_pp);
        const T uxz = T(0.25) * (zx_mm - zx_mp - zx_pm + zx_pp);
        const T uxxy = T(0.25) * (-xy_mm + xy_mp - xy_pm + xy_pp) + T(0.5) * (ym1 - yp1);
        const T uxyy = T(0.5) * (xm1 - xp1) + T(0.25) * (xy_pp - xy_mm - xy_mp + xy_pm);
        const T uyyz = T(0.25) * (-yz_mm + yz_mp - yz_pm + yz_pp) + T(0.5) * (zm1 - zp1);
        const T uyzz = T(0.5) * (ym1 - yp1) + T(0.25) * (yz_pp - yz_mm - yz_mp + yz_pm);
        const T uxxz = T(0.25) * (-zx_mm + zx_pm - zx_mp + zx_pp) + T(0.5) * (zm1 - zp1);
        const T uxzz = T(0.5) * (xm1 - xp1) + T(0.25) * (zx_pp - zx_mm - zx_pm + zx_mp);
        is4 += k7 * (uxy * uxy + uyz * uyz + uxz * uxz) +
               T(4.7) * (uxxy * uxxy + uxxz * uxxz + uxyy * uxyy + uxzz * uxzz + uyyz * uyyz + uyzz * uyzz);
        const T dxy1 = T(0.5) * (xy_pp - xy_pm) - cy, dxz1 = T(0.5) * (zx_pp - zx_mp) - cz;
        const T dxy2 = T(0.5) * (xy_mm - xy_mp) + cy, dxz2 = T(0.5) * (zx_mm - zx_pm) + cz;
        const T dxy3 = T(0.5) * (xy_pp - xy_mp) - cx, dyz3 = T(0.5) * (yz_pp - yz_pm) - cz;
        const T dxy4 = T(0.5) * (xy_mm - xy_pm) + cx, dyz4 = T(0.5) * (yz_mm - yz_mp) + cz;
        const T dyz5 = T(0.5) * (yz_pp - yz_mp) - cy, dxz5 = T(0.5) * (zx_pp - zx_pm) - cx;
        const T dyz6 = T(0.5) * (yz_mm - yz_pm) + cy, dxz6 = T(0.5) * (zx_mm - zx_mp) + cx;
        const T Cxy = k7 * uxy * uxy, Cyz = k7 * uyz * uyz, Cxz = k7 * uxz * uxz;
        is3[0] += Cxy + Cyz + Cxz;
        is3[1] += k7 * dxy1 * dxy1 + Cyz + k7 * dxz1 * dxz1;
        is3[2] += k7 * dxy2 * dxy2 + Cyz + k7 * dxz2 * dxz2;
        is3[3] += k7 * dxy3 * dxy3 + k7 * dyz3 * dyz3 + Cxz;
        is3[4] += k7 * dxy4 * dxy4 + k7 * dyz4 * dyz4 + Cxz;
        is3[5] += Cxy + k7 * dyz5 * dyz5 + k7 * dxz5 * dxz5;
        is3[6] += Cxy + k7 * dyz6 * dyz6 + k7 * dxz6 * dxz6;
    }
    asm volatile("" ::: "memory");
    {   // corners: u_xyz (reconstruction.c:108)
        const T uxyz = T(0.125) * (get(25) - get(26) - get(27) - get(28) + get(29) + get(30) + get(31) - get(32));
        is4 += T(61.0 / 48.0) * uxyz * uxyz;
    }

    // nonlinear weights, p = 4, eps = 1e-12 (reconstruction.c:118,205-218)
    const T eps = T(1.0e-12);
    const T g4 = T(p.g4), g3c = T(p.g3c), g3s = T(p.g3s);
    T w4;
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
        const T it = rcp64(total);
        wt = (w4 * it) * T(p.inv_g4);
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
}

// CWENO combination (reconstruction.c:223-257) for the face-relevant
// coefficients only, from the 13 axis values (u0 and the x, y, z arms
// m2, m1, p1, p2) and the weights.  lin = sum gamma_k R3_k, nl = sum w_k R3_k.
template <typename T>
__device__ __forceinline__ void cweno_combine(T u0, T xm2, T xm1, T xp1, T xp2, T ym2, T ym1, T yp1, T yp2, T zm2,
                                              T zm1, T zp1, T zp2, T wt, const T (&w3)[7], const CwArgs &p,
                                              Axis3<T> &o)
{
    const T g3s = T(p.g3s);
    const T gc = T(p.g3c) + T(4) * g3s;           // weight of stencils central along an axis
    T ux, uy, uz, uxx, uyy, uzz, uxxx, uyyy, uzzz;
    axis_linear(xm2, xm1, u0, xp1, xp2, ux, uxx, uxxx);
    axis_linear(ym2, ym1, u0, yp1, yp2, uy, uyy, uyyy);
    axis_linear(zm2, zm1, u0, zp1, zp2, uz, uzz, uzzz);
    const T cx = T(0.5) * (xp1 - xm1), cy = T(0.5) * (yp1 - ym1), cz = T(0.5) * (zp1 - zm1);
    const T dx1 = T(-1.5) * u0 + T(2) * xp1 - T(0.5) * xp2, dxx1 = T(0.5) * (u0 - T(2) * xp1 + xp2);
    const T dx2 = T(-2) * xm1 + T(0.5) * xm2 + T(1.5) * u0, dxx2 = T(0.5) * (xm2 - T(2) * xm1 + u0);
    const T dy3 = T(-1.5) * u0 + T(2) * yp1 - T(0.5) * yp2, dyy3 = T(0.5) * (u0 - T(2) * yp1 + yp2);
    const T dy4 = T(-2) * ym1 + T(0.5) * ym2 + T(1.5) * u0, dyy4 = T(0.5) * (ym2 - T(2) * ym1 + u0);
    const T dz5 = T(-1.5) * u0 + T(2) * zp1 - T(0.5) * zp2, dzz5 = T(0.5) * (u0 - T(2) * zp1 + zp2);
    const T dz6 = T(-2) * zm1 + T(0.5) * zm2 + T(1.5) * u0, dzz6 = T(0.5) * (zm2 - T(2) * zm1 + u0);
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
}

// Full CWENO4 from a register array of the 33 stencil values: returns the
// axis coefficients (c1,c2,c3 per axis) and wt_ratio; wn (if not null)
// receives the eight normalised weights (R4, R3[0..6]).
template <typename T>
__device__ __forceinline__ T cweno_axes(const T (&s)[33], const CwArgs &p, Axis3<T> &o, T *wn = nullptr)
{
    T wt, w3[7];
    cweno_weights<T>([&](int q) { return s[q]; }, p, wt, w3);
    if (wn) {
        wn[0] = wt * T(p.g4);
#pragma unroll
        for (int k = 0; k < 7; ++k) wn[1 + k] = w3[k];
    }
    cweno_combine<T>(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10], s[11], s[12], wt, w3, p, o);
    return wt;
}

// Full CWENO4 from values held in LDS, get(q) returning stencil slot q: the
// weights are formed group by group, then the 13 axis values are read again
// for the combination (the memory clobbers keep the compiler from holding all
// 33 values live across the weight arithmetic).
template <typename T, typename Get>
__device__ __forceinline__ void cweno_axes_lds(Get get, const CwArgs &p, Axis3<T> &o)
{
    T wt, w3[7];
    cweno_weights<T>(get, p, wt, w3);
    asm volatile("" ::: "memory");
    cweno_combine<T>(get(0), get(1), get(2), get(3), get(4), get(5), get(6), get(7), get(8), get(9), get(10),
                     get(11), get(12), wt, w3, p, o);
}

} // namespace acx
