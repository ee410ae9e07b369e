// acx_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Ghost cells are never materialised: a neighbour index outside [0, n) is
// mapped back into the owned range the way the reference's DMDA ghost fill
// does it, axis by axis (which composes the rules at edges and corners):
//   periodic     wrap (3d/weno4/plane_wave/hype.c:87-104);
//   adiabatic    mirror, component 0 negated (WAVE2,
//                3d/weno4/plane_wave/rhs_function.c:48-135);
//   reflective   mirror, normal velocity component 1+d negated (ACOUSTIC4,
//                3d/weno3/LinearisedEuler/rhs_function.c:47-176);
//   transmissive copy of the boundary cell (same file).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acx {

// Maps a (possibly ghost) index along one axis of boundary kind `kind`
// (ACX_PERIODIC 0, ACX_ADIABATIC 1, ACX_REFLECTIVE 2, ACX_TRANSMISSIVE 3).
// `sign` is multiplied by -1 for a mirror.  Indices further than 3 cells
// outside a non-periodic axis are only reached by threads whose results are
// discarded; they are clamped so that every load stays inside the array.
__device__ __forceinline__ int map_axis(int i, int n, int kind, float &sign)
{
    if (i >= 0 && i < n) return i;
    if (kind == 0) {
        while (i < 0) i += n;
        while (i >= n) i -= n;
        return i;
    }
    if (kind == 3) return i < 0 ? 0 : n - 1;
    sign = -sign;
    if (i < 0) {
        int r = -1 - i;
        return r < n ? r : n - 1;
    }
    int r = 2 * n - 1 - i;
    return r >= 0 ? r : 0;
}

// XCD-aware tile order.  Workgroups of a 1D launch are handed to the 8 XCDs
// round-robin (id % 8); each XCD has its own L2.  Give XCD x a contiguous run
// of logical tiles so that tiles sharing halo planes are resident on the
// same L2 at the same time.  Bijective for any workgroup count.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_logical_tile(int id, int nb)
{
    const int xcd = id % kNumXcd, slot = id / kNumXcd;
    const int base = nb / kNumXcd, rem = nb % kNumXcd;
    return xcd * base + (xcd < rem ? xcd : rem) + slot;
}

template <typename T> struct Vec2 { T a, b; };

// AoS cell access (4 components: p,u,v,w; or 2 components: u,v)
template <typename T, int NV> struct Cell { T v[NV]; };

template <typename T, int NV>
__device__ __forceinline__ Cell<T, NV> load_cell(const T *__restrict__ base, size_t idx)
{
    Cell<T, NV> c;
    if constexpr (sizeof(T) == 8 && NV == 4) {
        const double2 *p = reinterpret_cast<const double2 *>(base + idx * 4);
        double2 a = p[0], b = p[1];
        c.v[0] = a.x; c.v[1] = a.y; c.v[2] = b.x; c.v[3] = b.y;
    } else if constexpr (sizeof(T) == 8 && NV == 2) {
        const double2 *p = reinterpret_cast<const double2 *>(base + idx * 2);
        double2 a = p[0];
        c.v[0] = a.x; c.v[1] = a.y;
    } else if constexpr (sizeof(T) == 4 && NV == 4) {
        const float4 *p = reinterpret_cast<const float4 *>(base + idx * 4);
        float4 a = p[0];
        c.v[0] = a.x; c.v[1] = a.y; c.v[2] = a.z; c.v[3] = a.w;
    } else {
        const float2 *p = reinterpret_cast<const float2 *>(base + idx * 2);
        float2 a = p[0];
        c.v[0] = a.x; c.v[1] = a.y;
    }
    return c;
}

// Ghost signs of a mapped cell (sx, sy, sz: -1 where the x, y, z mapping
// mirrored): WAVE2 (NV = 2) negates component 0 once per mirror; ACOUSTIC4
// (NV = 4) negates the velocity component normal to each mirror.
template <typename T, int NV>
__device__ __forceinline__ void apply_bc_sign(Cell<T, NV> &c, float sx, float sy, float sz)
{
    if constexpr (NV == 4) {
        c.v[1] *= T(sx);
        c.v[2] *= T(sy);
        c.v[3] *= T(sz);
    } else {
        c.v[0] *= T(sx * sy * sz);
    }
}

template <typename T, int NV>
__device__ __forceinline__ void store_cell(T *__restrict__ base, size_t idx, const Cell<T, NV> &c)
{
    if constexpr (sizeof(T) == 8 && NV == 4) {
        double2 *p = reinterpret_cast<double2 *>(base + idx * 4);
        p[0] = make_double2(c.v[0], c.v[1]);
        p[1] = make_double2(c.v[2], c.v[3]);
    } else if constexpr (sizeof(T) == 8 && NV == 2) {
        double2 *p = reinterpret_cast<double2 *>(base + idx * 2);
        p[0] = make_double2(c.v[0], c.v[1]);
    } else if constexpr (sizeof(T) == 4 && NV == 4) {
        float4 *p = reinterpret_cast<float4 *>(base + idx * 4);
        p[0] = make_float4(c.v[0], c.v[1], c.v[2], c.v[3]);
    } else {
        float2 *p = reinterpret_cast<float2 *>(base + idx * 2);
        p[0] = make_float2(c.v[0], c.v[1]);
    }
}

// Non-finite watch.  finite_acc marks a thread whose output cell holds a NaN
// or +-Inf; watch_flush, called once after a tile's march by every thread of
// the block, lets the first marked lane of each wavefront store 1 to *nf, a
// host-mapped word (no store on the finite path).
template <typename T, int NV>
__device__ __forceinline__ void finite_acc(bool &bad, const Cell<T, NV> &c)
{
#pragma unroll
    for (int v = 0; v < NV; ++v) bad |= !__builtin_isfinite(c.v[v]);
}

__device__ __forceinline__ void watch_flush(unsigned *nf, bool bad)
{
    if (!nf) return;
    const unsigned long long m = __ballot(bad);
    if (bad && (int)__lane_id() == __ffsll((long long)m) - 1) *(volatile unsigned *)nf = 1u;
}

// Streaming forms of load_cell / store_cell (nontemporal hint: the line is
// not kept in L2 for reuse), for data that no other tile of the launch reads.
typedef double acx_dbl2 __attribute__((ext_vector_type(2)));
typedef float acx_flt4 __attribute__((ext_vector_type(4)));

template <typename T, int NV>
__device__ __forceinline__ Cell<T, NV> load_cell_nt(const T *__restrict__ base, size_t idx)
{
    Cell<T, NV> c;
    if constexpr (sizeof(T) == 8 && NV == 4) {
        const acx_dbl2 *p = reinterpret_cast<const acx_dbl2 *>(base + idx * 4);
        const acx_dbl2 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
        c.v[0] = a.x; c.v[1] = a.y; c.v[2] = b.x; c.v[3] = b.y;
    } else if constexpr (sizeof(T) == 4 && NV == 4) {
        const acx_flt4 a = __builtin_nontemporal_load(reinterpret_cast<const acx_flt4 *>(base + idx * 4));
        c.v[0] = a.x; c.v[1] = a.y; c.v[2] = a.z; c.v[3] = a.w;
    } else {
        c = load_cell<T, NV>(base, idx);
    }
    return c;
}

template <typename T, int NV>
__device__ __forceinline__ void store_cell_nt(T *__restrict__ base, size_t idx, const Cell<T, NV> &c)
{
    if constexpr (sizeof(T) == 8 && NV == 4) {
        acx_dbl2 *p = reinterpret_cast<acx_dbl2 *>(base + idx * 4);
        __builtin_nontemporal_store(acx_dbl2{c.v[0], c.v[1]}, p);
        __builtin_nontemporal_store(acx_dbl2{c.v[2], c.v[3]}, p + 1);
    } else if constexpr (sizeof(T) == 4 && NV == 4) {
        __builtin_nontemporal_store(acx_flt4{c.v[0], c.v[1], c.v[2], c.v[3]}, reinterpret_cast<acx_flt4 *>(base + idx * 4));
    } else {
        store_cell<T, NV>(base, idx, c);
    }
}


// Exchanges a's lanes 32-63 with b's lanes 0-31 (v_permlane32_swap, two
// dwords per double): afterwards a holds [a of lanes 0-31, b of lanes 0-31]
// and b holds [a of lanes 32-63, b of lanes 32-63].  Every lane of the
// wavefront must execute it.
__device__ __forceinline__ void swap_halves32(double &a, double &b)
{
    const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// Linear-scheme face stencils folded into 7-point cell stencils.
// With the shipped weights (gamma_hi = 1, gamma_lo = 0) the CWENO4 face
// averages are q_L(i+1/2) = (q_{i-2} - 6q_{i-1} + 20q_i + 10q_{i+1} - q_{i+2})/24
// and the mirror image q_R; the flux-difference of the LLF / central fluxes
// then needs only
//   A7 = 24 * [ (q_L+q_R)(i+1/2) - (q_L+q_R)(i-1/2) ]
//      = 37(q1-q-1) - 8(q2-q-2) + (q3-q-3)
//   B7 = 24 * [ (q_R-q_L)(i+1/2) - (q_R-q_L)(i-1/2) ]
//      = (q3+q-3) - 6(q2+q-2) + 15(q1+q-1) - 20 q0
// and for the central flux of face-averaged normal gradients (WAVE2)
//   L7 = 48 h^2 * div = -(q3+q-3) + 2(q2+q-2) + 49(q1+q-1) - 100 q0.
template <typename T>
__device__ __forceinline__ T stencil_A7(T m3, T m2, T m1, T p1, T p2, T p3)
{
    return T(37) * (p1 - m1) - T(8) * (p2 - m2) + (p3 - m3);
}
template <typename T>
__device__ __forceinline__ T stencil_B7(T m3, T m2, T m1, T q0, T p1, T p2, T p3)
{
    return (p3 + m3) - T(6) * (p2 + m2) + T(15) * (p1 + m1) - T(20) * q0;
}
template <typename T>
__device__ __forceinline__ T stencil_L7(T m3, T m2, T m1, T q0, T p1, T p2, T p3)
{
    return T(2) * (p2 + m2) - (p3 + m3) + T(49) * (p1 + m1) - T(100) * q0;
}

// Error-free sum of two floats (Knuth's TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

} // namespace acx
