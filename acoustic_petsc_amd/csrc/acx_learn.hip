// acx_learn.hip -- learned troubled-cell indicator (include/acx_learn.h):
// features, labels from the solver's own CWENO weights, MFMA forward,
// fp32 backward + Adam, fused grid inference, flag dilation.
//
// MFMA: v_mfma_f32_16x16x32_bf16 (gfx950).  Per wavefront 16 samples are the
// N columns; lane l holds sample n = l % 16 and the K slice 8*(l/16)..+7 of
// the B operand; the A operand (weights) holds row l % 16 and the same K
// slice; the accumulator holds rows 4*(l/16)..+3 of column l % 16.  The two
// 16-row halves of a 32-wide layer therefore leave lane l with the outputs
// {4g..4g+3, 16+4g..16+4g+3} (g = l/16), which is used as is for the next
// layer's K slice by permuting that layer's weight columns once at packing.
#include <math.h>
#include <string.h>
#include "acx_device.h"
#include "acx_internal.h"
#include "acx_cweno_core.h"
#include "../../include/acx_learn.h"

namespace acx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct MlpPacked {
    bf16x8 w1[2][64];   // layer 1 A operand per 16-row half and lane
    bf16x8 w2[2][64];   // layer 2 A operand, K permuted to the layer-1 output order
    float b1[32], b2[32], w3[32], b3;
    float act;          // hidden activation: 0 ReLU, 1 tanh
};

// output feature held in slot j (0..7) of a lane of group g
__host__ __device__ inline int kperm(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

// Hidden activations.  ReLU (the default): one VALU max per unit, max(x, 0)
// as one integer max on the bit pattern (negative floats and -0 have
// negative int patterns; fmaxf would add a NaN-quieting step).  tanh:
// 1 - 2 / (exp(2x) + 1), an exponential and a reciprocal per unit, which
// makes the fused inference pass bound by the quarter-rate transcendental
// issue (64 units per sample); exp overflow gives 1, underflow -1.
__device__ __forceinline__ float act_relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
__device__ __forceinline__ float act_tanh(float x) { return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f); }
__device__ __forceinline__ float act(float x, bool th) { return th ? act_tanh(x) : act_relu(x); }

// ---- features --------------------------------------------------------------

// a[axis][k]: value at offset k-2 along the axis.  18 normalised undivided
// differences (per axis: 4 first differences, 2 second differences) and a
// log-amplitude term that tells the network how the differences compare
// with the absolute eps = 1e-12 of the smoothness indicators.
__device__ __forceinline__ void features19(const float (&a)[3][5], float (&f)[ACX_MLP_NFEAT])
{
    float d[3][4];
    float S = 0.f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            d[ax][k] = a[ax][k + 1] - a[ax][k];
            S = fmaxf(S, fabsf(d[ax][k]));
        }
    const float inv = S > 0.f ? __builtin_amdgcn_rcpf(S) : 0.f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
        for (int k = 0; k < 4; ++k) f[6 * ax + k] = d[ax][k] * inv;
        f[6 * ax + 4] = (d[ax][1] - d[ax][0]) * inv;
        f[6 * ax + 5] = (d[ax][3] - d[ax][2]) * inv;
    }
    // log10(S) = log2(S) * log10(2)
    f[18] = fminf(fmaxf((__log2f(S + 1e-30f) * 0.30102999566f + 6.f) * (1.f / 6.f), -1.f), 2.f);
}

// The 5-point arms of cell (i, j, k), component v, on a periodic grid:
// a[axis][q] = value at offset q - 2 (32-bit index math; grids stay below
// 2^32 / 4 values).
__device__ __forceinline__ void gather_arms(const float *__restrict__ U, const float *__restrict__ D, int nx, int ny,
                                            int nz, int i, int j, int k, int v, float (&a)[3][5])
{
    auto W = [](int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); };
    const unsigned sx = 4u, sy = 4u * (unsigned)nx, sz = 4u * (unsigned)nx * (unsigned)ny;
    const unsigned base = (unsigned)k * sz + (unsigned)j * sy + (unsigned)i * sx + (unsigned)v;
    auto at = [&](unsigned idx) { return D ? U[idx] + D[idx] : U[idx]; };
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
        a[0][o + 2] = at(base + (unsigned)(W(i + o, nx) - i) * sx);
        a[1][o + 2] = at(base + (unsigned)(W(j + o, ny) - j) * sy);
        a[2][o + 2] = at(base + (unsigned)(W(k + o, nz) - k) * sz);
    }
}

__global__ void __launch_bounds__(256) k_features(const float *__restrict__ U, int nx, int ny, int nz, float *__restrict__ X)
{
    const size_t ns = (size_t)nx * ny * nz * 4;
    const size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const size_t cell = s / 4;
    const int v = (int)(s % 4);
    const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / ((size_t)nx * ny));
    float a[3][5], f[ACX_MLP_NFEAT];
    gather_arms(U, nullptr, nx, ny, nz, i, j, k, v, a);
    features19(a, f);
    float *o = X + s * ACX_MLP_IN;
#pragma unroll
    for (int q = 0; q < ACX_MLP_IN; ++q) o[q] = q < ACX_MLP_NFEAT ? f[q] : 0.f;
}

// ---- labels from the solver's weights -------------------------------------------

__global__ void __launch_bounds__(256) k_labels(const double *__restrict__ U, CwArgs p, double thr,
                                                uint8_t *__restrict__ labels, float *__restrict__ wt)
{
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t cell = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / ((size_t)nx * ny));
    auto W = [](int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); };
    const int dx[33] = {0, -2, -1, 1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, 1, -1, -1, 1, -1};
    const int dy[33] = {0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, 1, -1, 1, -1, -1};
    const int dz[33] = {0, 0, 0, 0, 0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, 1, -1, 1, -1, -1, -1};
    for (int v = 0; v < 4; ++v) {
        double st[33];
#pragma unroll
        for (int q = 0; q < 33; ++q)
            st[q] = U[(((size_t)W(k + dz[q], nz) * ny + W(j + dy[q], ny)) * nx + W(i + dx[q], nx)) * 4 + v];
        Axis3<double> ax;
        double wn[8];
        cweno_axes<double>(st, p, ax, wn);
        double dev = fabs(wn[0] - p.g4) + fabs(wn[1] - p.g3c);
#pragma unroll
        for (int q = 2; q < 8; ++q) dev += fabs(wn[q] - p.g3s);
        labels[cell * 4 + v] = dev > thr ? 1 : 0;
        if (wt) wt[cell * 4 + v] = (float)dev;
    }
}

// ---- full-WENO judge of the selective step ---------------------------------------
//
// The full CWENO4 reconstruction (FP32, ratio-form weights, grid gammas) of
// every cell and component of the FP32 state the selective step advances.
// For the (cell, component) pairs a window of selective steps ran with the
// linear polynomial (bit clear in `used`) the judge measures what that
// choice changed: the largest difference between the CWENO4 and the linear
// face averages over the six faces, |dc2|/6 + |dc1/2 + dc3/20| along each
// axis (the face-average collapse of the basis).  A difference above `tol`
// is a miss (counted); the largest difference is kept (float bits, atomic
// max of non-negative values).  Independently, label bit v of a cell is set
// when the eight normalised weights of component v leave the linear weights
// by more than `thr` (the training label), and out = base | labels.
// One block judges a JT_BX x JT_BY column tile over JT_ZC planes: the
// stencil region of the current plane and of the two planes on either side
// (2-cell x/y halo) sits in a 5-plane LDS ring, one array per component, so
// the 33 values of every cell are read from LDS; plane z+3 is loaded while
// plane z is judged.
constexpr int JT_BX = 32, JT_BY = 8, JT_ZC = 16;
constexpr int JT_LX = JT_BX + 4, JT_LY = JT_BY + 4;

__global__ void __launch_bounds__(JT_BX *JT_BY) k_judge(const float *__restrict__ HI, CwArgs p, float thr, float tol,
                                                          const uint8_t *__restrict__ used,
                                                          const uint8_t *__restrict__ base, uint8_t *__restrict__ out,
                                                          unsigned long long *__restrict__ missed,
                                                          unsigned *__restrict__ worst)
{
    __shared__ float ring[5][4][JT_LY][JT_LX];
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int ntx = (nx + JT_BX - 1) / JT_BX, nty = (ny + JT_BY - 1) / JT_BY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * JT_BX, y0 = ((b / ntx) % nty) * JT_BY, zb = (b / (ntx * nty)) * JT_ZC;
    const int ze = min(zb + JT_ZC, nz);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * JT_BX + tx;
    const bool active = x0 + tx < nx && y0 + ty < ny;
    const size_t plane = (size_t)nx * ny;
    auto wrap = [](int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); };
    // the stencil region of plane zz into ring slot sl (float4 loads, one cell per thread and pass)
    auto load_plane = [&](int zz, int sl) {
        const size_t zo = (size_t)wrap(zz, nz) * plane;
        for (int c = tid; c < JT_LX * JT_LY; c += JT_BX * JT_BY) {
            const int ly = c / JT_LX, lx = c - ly * JT_LX;
            const int gx = wrap(x0 + lx - 2, nx), gy = wrap(y0 + ly - 2, ny);
            const float4 v = *reinterpret_cast<const float4 *>(HI + (zo + (size_t)gy * nx + gx) * 4);
            ring[sl][0][ly][lx] = v.x;
            ring[sl][1][ly][lx] = v.y;
            ring[sl][2][ly][lx] = v.z;
            ring[sl][3][ly][lx] = v.w;
        }
    };
    // stencil slots of cweno_weights: centre, x/y/z arms (m2 m1 p1 p2), xy/yz/zx diagonals, corners
    constexpr int DX[33] = {0, -2, -1, 1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, 1, -1, -1, 1, -1};
    constexpr int DY[33] = {0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, 1, -1, 1, -1, -1};
    constexpr int DZ[33] = {0, 0, 0, 0, 0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, 1, -1, 1, -1, -1, -1};
    for (int d = -2; d <= 2; ++d) load_plane(zb + d, d + 2);
    unsigned miss = 0;
    float wmax = 0.f;
    for (int z = zb; z < ze; ++z) {
        __syncthreads();
        const int s0 = (z - zb + 2) % 5;   // slot of plane z; plane z+d is in slot (s0 + d) mod 5
        if (active) {
            const size_t cell = (size_t)z * plane + (size_t)(y0 + ty) * nx + (x0 + tx);
            const unsigned uflag = used ? used[cell] : 0xFu;
            unsigned lab = 0;
            for (int v = 0; v < 4; ++v) {
                auto get = [&](int q) {
                    return ring[(s0 + 5 + DZ[q]) % 5][v][ty + 2 + DY[q]][tx + 2 + DX[q]];
                };
                float wt, w3[7];
                cweno_weights<float>(get, p, wt, w3);
                float dev = fabsf(wt * float(p.g4) - float(p.g4)) + fabsf(w3[0] - float(p.g3c));
#pragma unroll
                for (int q = 1; q < 7; ++q) dev += fabsf(w3[q] - float(p.g3s));
                if (dev > thr) lab |= 1u << v;
                if (!((uflag >> v) & 1u)) {
                    // CWENO4 and linear face averages of the cell, largest difference over the six faces
                    Axis3<float> o;
                    cweno_combine<float>(get(0), get(1), get(2), get(3), get(4), get(5), get(6), get(7), get(8),
                                         get(9), get(10), get(11), get(12), wt, w3, p, o);
                    float d = 0.f;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        float c1, c2, c3;
                        const int m2 = 1 + 4 * a;   // arm slots m2 m1 p1 p2 of axis a
                        axis_linear(get(m2), get(m2 + 1), get(0), get(m2 + 2), get(m2 + 3), c1, c2, c3);
                        d = fmaxf(d, fabsf(o.c2[a] - c2) * (1.f / 6.f) +
                                         fabsf(0.5f * (o.c1[a] - c1) + 0.05f * (o.c3[a] - c3)));
                    }
                    wmax = fmaxf(wmax, d);
                    if (d > tol) ++miss;
                }
            }
            if (out) out[cell] = (uint8_t)((base ? base[cell] : 0) | lab);
        }
        __syncthreads();   // every thread is done with the slot of plane z-2
        if (z + 1 < ze) load_plane(z + 3, (s0 + 3) % 5);
    }
    if (missed) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            miss += __shfl_down(miss, o, 64);
            wmax = fmaxf(wmax, __shfl_down(wmax, o, 64));
        }
        if ((tid & 63) == 0) {
            if (miss) atomicAdd(missed, (unsigned long long)miss);
            if (wmax > 0.f) atomicMax(worst, __float_as_uint(wmax));   // non-negative floats order as their bits
        }
    }
}

// ---- MLP forward on MFMA ------------------------------------------------------

// Logit z of 16 samples per wavefront (y = sigmoid(z)); the biases enter as
// the initial accumulators of the MFMAs (accumulator row 4g + i of lane l).
__device__ __forceinline__ float mlp_wave(const MlpPacked &P, const bf16x8 &x, int lane, float (*h1out)[8],
                                          float (*h2out)[8])
{
    const int g = lane >> 4;
    floatx4 i0, i1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        i0[i] = P.b1[4 * g + i];
        i1[i] = P.b1[16 + 4 * g + i];
    }
    const floatx4 a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.w1[0][lane], x, i0, 0, 0, 0);
    const floatx4 a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.w1[1][lane], x, i1, 0, 0, 0);
    const bool th = P.act != 0.f;
    bf16x8 hb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        hb[i] = (__bf16)act(a0[i], th);
        hb[4 + i] = (__bf16)act(a1[i], th);
    }
    if (h1out) {
#pragma unroll
        for (int q = 0; q < 8; ++q) (*h1out)[q] = (float)hb[q];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        i0[i] = P.b2[4 * g + i];
        i1[i] = P.b2[16 + 4 * g + i];
    }
    const floatx4 c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.w2[0][lane], hb, i0, 0, 0, 0);
    const floatx4 c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.w2[1][lane], hb, i1, 0, 0, 0);
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float u0 = act(c0[i], th), u1 = act(c1[i], th);
        if (h2out) { (*h2out)[i] = u0; (*h2out)[4 + i] = u1; }
        acc += P.w3[4 * g + i] * u0 + P.w3[16 + 4 * g + i] * u1;
    }
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    return acc + P.b3;
}

__global__ void __launch_bounds__(256) k_mlp_forward(const MlpPacked *__restrict__ Pg, const float *__restrict__ X,
                                                     int n, float *__restrict__ Y, float *__restrict__ H1,
                                                     float *__restrict__ H2)
{
    __shared__ MlpPacked P;
    for (int q = threadIdx.x; q < (int)(sizeof(MlpPacked) / 4); q += blockDim.x)
        reinterpret_cast<float *>(&P)[q] = reinterpret_cast<const float *>(Pg)[q];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, sl = lane & 15;
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int base = (blockIdx.x * (blockDim.x >> 6) + wv) * 16; base < n; base += nw * 16) {
        const int s = base + sl;
        bf16x8 x;
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = (__bf16)(s < n ? X[(size_t)s * ACX_MLP_IN + 8 * g + q] : 0.f);
        float h1[8], h2[8];
        const float y = 1.f / (1.f + __expf(-mlp_wave(P, x, lane, H1 ? &h1 : nullptr, H2 ? &h2 : nullptr)));
        if (s < n) {
            if (g == 0) Y[s] = y;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (H1) H1[(size_t)s * 32 + kperm(g, q)] = h1[q];
                if (H2) H2[(size_t)s * 32 + kperm(g, q)] = h2[q];
            }
        }
    }
}

// ---- fused grid inference -------------------------------------------------------

constexpr int IND_WAVES = 4;

// Each wave pass covers 16 cells of an x-row x 4 components = 64 samples:
// lane l forms the features of sample l (cell i0 + l/4 of the row, component
// l%4; the 64 lanes read 64 consecutive floats per stencil point), parks them
// in LDS, and the wave then runs four 16-sample MFMA batches over them.
__global__ void __launch_bounds__(64 * IND_WAVES) k_indicator(const MlpPacked *__restrict__ Pg,
                                                             const float *__restrict__ HI,
                                                             const float *__restrict__ D, int nx, int ny, int nz,
                                                             float zthr, uint8_t *__restrict__ flags)
{
    __shared__ MlpPacked P;
    // features parked as the bf16 operands the MFMA takes (rows of 40: the
    // 16-byte K slices stay aligned); half the LDS of fp32 rows, so more
    // waves fit per CU
    constexpr int FROW = ACX_MLP_IN + 8;
    __shared__ __align__(16) __bf16 feat[IND_WAVES][64][FROW];
    for (int q = threadIdx.x; q < (int)(sizeof(MlpPacked) / 4); q += blockDim.x)
        reinterpret_cast<float *>(&P)[q] = reinterpret_cast<const float *>(Pg)[q];
    __syncthreads();
    // one wave pass = 16 cells of one x-row (segment seg of row (j, k)) x 4
    // components; the pass index and its row are wave-uniform (scalar math)
    const int lane = threadIdx.x & 63, g = lane >> 4, sl = lane & 15;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    __bf16(*F)[FROW] = feat[wv];
    const unsigned segs = (unsigned)(nx + 15) / 16u;
    const unsigned npass = segs * (unsigned)ny * (unsigned)nz;
    const unsigned nw = gridDim.x * IND_WAVES;
    for (unsigned ps = blockIdx.x * IND_WAVES + (unsigned)wv; ps < npass; ps += nw) {
        const unsigned row = ps / segs;
        const int i0 = (int)(ps - row * segs) * 16;
        const int j = (int)(row % (unsigned)ny), k = (int)(row / (unsigned)ny);
        {
            const int i = i0 + (lane >> 2);
            const int v = lane & 3;
            float f[ACX_MLP_NFEAT];
            if (i < nx) {
                float a[3][5];
                gather_arms(HI, D, nx, ny, nz, i, j, k, v, a);
                features19(a, f);
            } else {
#pragma unroll
                for (int q = 0; q < ACX_MLP_NFEAT; ++q) f[q] = 0.f;
            }
            // four 16-byte stores per row (a 2-byte store per feature put four
            // lanes on each LDS bank: rows are 80 bytes apart)
            bf16x8 c[ACX_MLP_IN / 8];
#pragma unroll
            for (int q = 0; q < ACX_MLP_IN; ++q) c[q / 8][q % 8] = (__bf16)(q < ACX_MLP_NFEAT ? f[q] : 0.f);
#pragma unroll
            for (int h = 0; h < ACX_MLP_IN / 8; ++h) *reinterpret_cast<bf16x8 *>(&F[lane][8 * h]) = c[h];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const bf16x8 x = *reinterpret_cast<const bf16x8 *>(&F[16 * b + sl][8 * g]);
            const float z = mlp_wave(P, x, lane, nullptr, nullptr);
            const unsigned long long m = __ballot(z > zthr);   // bit n (< 16): sample 16b + n
            const int ic = i0 + 4 * b + lane;
            if (lane < 4 && ic < nx) flags[(size_t)row * nx + ic] = (uint8_t)((m >> (4 * lane)) & 0xF);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();   // every read of F is done before the next pass writes it
    }
}

// ---- backward (fp32) and Adam ------------------------------------------------------

constexpr int BW_SAMPLES = 64;

__global__ void __launch_bounds__(256) k_mlp_backward(const float *__restrict__ prm, const float *__restrict__ X,
                                                      const float *__restrict__ H1, const float *__restrict__ H2,
                                                      const float *__restrict__ Y, const uint8_t *__restrict__ T,
                                                      int n, float pos_weight, float *__restrict__ grad, int tanh_act)
{
    // derivative of the activation from its output h: ReLU 1 if h > 0, tanh 1 - h^2
    auto dact = [&](float h) { return tanh_act ? 1.f - h * h : (h > 0.f ? 1.f : 0.f); };
    const float *W1 = prm, *W2 = prm + 1056, *w3 = prm + 2112;
    float *gW1 = grad, *gb1 = grad + 1024, *gW2 = grad + 1056, *gb2 = grad + 2080, *gw3 = grad + 2112, *gb3 = grad + 2144;
    __shared__ float sW2[32][33];
    __shared__ float dh2[BW_SAMPLES][33], h1s[BW_SAMPLES][33], dh1[BW_SAMPLES][33], xs[BW_SAMPLES][33];
    __shared__ float dzs[BW_SAMPLES], h2s[BW_SAMPLES][33];
    for (int q = threadIdx.x; q < 1024; q += blockDim.x) sW2[q / 32][q % 32] = bf16_round(W2[q]);
    const int s0 = blockIdx.x * BW_SAMPLES;
    for (int q = threadIdx.x; q < BW_SAMPLES * 32; q += blockDim.x) {
        const int a = q / 32, b = q % 32, s = s0 + a;
        const bool ok = s < n;
        h1s[a][b] = ok ? H1[(size_t)s * 32 + b] : 0.f;
        h2s[a][b] = ok ? H2[(size_t)s * 32 + b] : 0.f;
        xs[a][b] = ok ? bf16_round(X[(size_t)s * ACX_MLP_IN + b]) : 0.f;
    }
    __syncthreads();
    if (threadIdx.x < BW_SAMPLES) {
        const int a = threadIdx.x, s = s0 + a;
        float dz = 0.f;
        if (s < n) {
            const float t = T[s] ? 1.f : 0.f, c = T[s] ? pos_weight : 1.f;
            dz = c * (Y[s] - t) / (float)n;
        }
        dzs[a] = dz;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < BW_SAMPLES * 32; q += blockDim.x) {
        const int a = q / 32, j = q % 32;
        dh2[a][j] = dzs[a] * w3[j] * dact(h2s[a][j]);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < BW_SAMPLES * 32; q += blockDim.x) {
        const int a = q / 32, kk = q % 32;
        float acc = 0.f;
#pragma unroll 8
        for (int j = 0; j < 32; ++j) acc += dh2[a][j] * sW2[j][kk];
        dh1[a][kk] = acc * dact(h1s[a][kk]);
    }
    __syncthreads();
    // parameter gradients: sum over the block's samples, one atomic per entry
    for (int q = threadIdx.x; q < 1024; q += blockDim.x) {
        const int r = q / 32, c = q % 32;
        float a1 = 0.f, a2 = 0.f;
        for (int a = 0; a < BW_SAMPLES; ++a) {
            a1 += dh1[a][r] * xs[a][c];
            a2 += dh2[a][r] * h1s[a][c];
        }
        atomicAdd(&gW1[q], a1);
        atomicAdd(&gW2[q], a2);
    }
    if (threadIdx.x < 32) {
        const int r = threadIdx.x;
        float b1 = 0.f, b2 = 0.f, w = 0.f;
        for (int a = 0; a < BW_SAMPLES; ++a) {
            b1 += dh1[a][r];
            b2 += dh2[a][r];
            w += dzs[a] * h2s[a][r];
        }
        atomicAdd(&gb1[r], b1);
        atomicAdd(&gb2[r], b2);
        atomicAdd(&gw3[r], w);
    }
    if (threadIdx.x == 32) {
        float b = 0.f;
        for (int a = 0; a < BW_SAMPLES; ++a) b += dzs[a];
        atomicAdd(gb3, b);
    }
    (void)W1;
}

__global__ void k_fill(float *p, int n, float v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_adam(float *__restrict__ prm, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_pack(const float *__restrict__ prm, MlpPacked *__restrict__ P, float act)
{
    const float *W1 = prm, *b1 = prm + 1024, *W2 = prm + 1056, *b2 = prm + 2080, *w3 = prm + 2112, *b3 = prm + 2144;
    const int t = threadIdx.x;   // 128 threads: half b = t / 64, lane = t % 64
    const int b = t >> 6, lane = t & 63, g = lane >> 4, row = 16 * b + (lane & 15);
    bf16x8 a1, a2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a1[j] = (__bf16)W1[row * 32 + 8 * g + j];
        a2[j] = (__bf16)W2[row * 32 + kperm(g, j)];
    }
    P->w1[b][lane] = a1;
    P->w2[b][lane] = a2;
    if (t < 32) {
        P->b1[t] = b1[t];
        P->b2[t] = b2[t];
        P->w3[t] = w3[t];
    }
    if (t == 0) {
        P->b3 = b3[0];
        P->act = act;
    }
}

// Block = 64 x 4 columns marching a chunk of DIL_ZC planes; the own column
// of planes k-1, k, k+1 stays in registers (no per-cell index division).
constexpr int DIL_BX = 64, DIL_BY = 4, DIL_ZC = 16;

__global__ void __launch_bounds__(DIL_BX *DIL_BY) k_dilate(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                          int nx, int ny, int nz)
{
    const int i = blockIdx.x * DIL_BX + threadIdx.x, j = blockIdx.y * DIL_BY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const size_t plane = (size_t)nx * ny;
    const int z0 = blockIdx.z * DIL_ZC, z1 = min(z0 + DIL_ZC, nz);
    auto at = [&](int k, int y, int x) { return in[(size_t)k * plane + (size_t)y * nx + x]; };
    uint8_t below = at(z0 == 0 ? nz - 1 : z0 - 1, j, i), cur = at(z0, j, i);
    for (int k = z0; k < z1; ++k) {
        const uint8_t above = at(k == nz - 1 ? 0 : k + 1, j, i);
        const uint8_t any = below | cur | above | at(k, j, im) | at(k, j, ip) | at(k, jm, i) | at(k, jp, i);
        out[(size_t)k * plane + (size_t)j * nx + i] = any ? 0xF : 0;   // the four acoustic components move together
        below = cur;
        cur = above;
    }
}

__global__ void k_count(const uint8_t *__restrict__ f, size_t n, unsigned long long *__restrict__ out)
{
    unsigned long long acc = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += __popc((unsigned)f[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

static int check_ac4(const acx_grid *g)
{
    if (!g || g->system != ACX_ACOUSTIC4) return set_error(ACX_EINVAL, "learned indicator: ACOUSTIC4 grid required");
    for (int d = 0; d < 3; ++d) {
        if (g->n[d] < 5) return set_error(ACX_EINVAL, "learned indicator: at least 5 cells per axis");
        if (g->bc[d] != ACX_PERIODIC) return set_error(ACX_EUNSUPPORTED, "learned indicator: periodic grids only");
    }
    if ((size_t)g->n[0] * g->n[1] * g->n[2] * 4 > 0xffffffffull)
        return set_error(ACX_EUNSUPPORTED, "learned indicator: more than 2^30 cells");
    return ACX_OK;
}

} // namespace acx

using namespace acx;

extern "C" {

size_t acx_mlp_packed_bytes(void) { return sizeof(MlpPacked); }

int acx_mlp_pack_act(const float *params, int activation, void *packed, void *stream)
{
    if (!params || !packed) return set_error(ACX_EINVAL, "mlp pack: null pointer");
    if (activation != ACX_MLP_RELU && activation != ACX_MLP_TANH) return set_error(ACX_EINVAL, "mlp pack: activation");
    hipLaunchKernelGGL(k_pack, dim3(1), dim3(128), 0, (hipStream_t)stream, params, (MlpPacked *)packed,
                       activation == ACX_MLP_TANH ? 1.f : 0.f);
    return hip_check(hipGetLastError(), "mlp pack");
}

int acx_mlp_pack(const float *params, void *packed, void *stream)
{
    return acx_mlp_pack_act(params, ACX_MLP_RELU, packed, stream);
}

int acx_features_f32(const acx_grid *g, const float *U, float *X, void *stream)
{
    int rc = check_ac4(g);
    if (rc) return rc;
    if (!U || !X) return set_error(ACX_EINVAL, "features: null pointer");
    const size_t ns = (size_t)g->n[0] * g->n[1] * g->n[2] * 4;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, (hipStream_t)stream, U, g->n[0],
                       g->n[1], g->n[2], X);
    return hip_check(hipGetLastError(), "features");
}

int acx_weno_labels_f64(const acx_grid *g, const double *U, double thr, uint8_t *labels, float *wt, void *stream)
{
    int rc = check_ac4(g);
    if (rc) return rc;
    if (!U || !labels) return set_error(ACX_EINVAL, "labels: null pointer");
    CwArgs p;
    memset(&p, 0, sizeof(p));
    p.nx = g->n[0]; p.ny = g->n[1]; p.nz = g->n[2];
    p.g4 = g->gamma_hi;
    p.g3c = (1.0 - g->gamma_hi) * g->gamma_lo;
    p.g3s = (1. / 6.) * (1 - g->gamma_hi) * (1 - g->gamma_lo);
    p.inv_g4 = 1.0 / g->gamma_hi;
    const size_t nc = (size_t)g->n[0] * g->n[1] * g->n[2];
    hipLaunchKernelGGL(k_labels, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, (hipStream_t)stream, U, p, thr,
                       labels, wt);
    return hip_check(hipGetLastError(), "labels");
}

int acx_mlp_forward(const void *packed, const float *X, int n, float *Y, float *H1, float *H2, void *stream)
{
    if (!packed || !X || !Y || n < 0) return set_error(ACX_EINVAL, "mlp forward: bad arguments");
    if (n == 0) return ACX_OK;
    const int waves = (n + 15) / 16;
    const int blocks = waves / 4 + 1 < 4096 ? waves / 4 + 1 : 4096;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const MlpPacked *)packed, X, n, Y,
                       H1, H2);
    return hip_check(hipGetLastError(), "mlp forward");
}

int acx_mlp_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                     const uint8_t *T, int n, float pos_weight, float *grad, void *stream)
{
    return acx_mlp_backward_act(params, ACX_MLP_RELU, X, H1, H2, Y, T, n, pos_weight, grad, stream);
}

int acx_mlp_backward_act(const float *params, int activation, const float *X, const float *H1, const float *H2,
                         const float *Y, const uint8_t *T, int n, float pos_weight, float *grad, void *stream)
{
    if (!params || !X || !H1 || !H2 || !Y || !T || !grad || n <= 0) return set_error(ACX_EINVAL, "mlp backward: bad arguments");
    if (activation != ACX_MLP_RELU && activation != ACX_MLP_TANH) return set_error(ACX_EINVAL, "mlp backward: activation");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill, dim3((ACX_MLP_NPARAM + 255) / 256), dim3(256), 0, s, grad, ACX_MLP_NPARAM, 0.f);
    hipLaunchKernelGGL(k_mlp_backward, dim3((n + BW_SAMPLES - 1) / BW_SAMPLES), dim3(256), 0, s, params, X, H1, H2, Y,
                       T, n, pos_weight, grad, activation == ACX_MLP_TANH ? 1 : 0);
    return hip_check(hipGetLastError(), "mlp backward");
}

int acx_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, float beta1, float beta2,
                 float eps, void *stream)
{
    if (!params || !grad || !m || !v || step < 1) return set_error(ACX_EINVAL, "adam: bad arguments");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((ACX_MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, grad, m, v,
                       ACX_MLP_NPARAM, lr, beta1, beta2, eps, c1, c2);
    return hip_check(hipGetLastError(), "adam");
}

int acx_indicator_flags(const acx_grid *g, const float *HI, const float *D, const void *packed, float thr,
                        uint8_t *flags, void *stream)
{
    int rc = check_ac4(g);
    if (rc) return rc;
    if (!HI || !packed || !flags) return set_error(ACX_EINVAL, "indicator: null pointer");
    // sigmoid(z) > thr  <=>  z > log(thr / (1 - thr)) for 0 < thr < 1
    const float zthr = !(thr > 0.f) ? -INFINITY : (thr >= 1.f ? INFINITY : logf(thr / (1.f - thr)));
    const size_t groups = (size_t)((g->n[0] + 15) / 16) * g->n[1] * g->n[2];   // one wave pass per 16 cells of a row
    size_t blocks = (groups + IND_WAVES - 1) / IND_WAVES;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_indicator, dim3((unsigned)blocks), dim3(64 * IND_WAVES), 0, (hipStream_t)stream,
                       (const MlpPacked *)packed,
                       HI, D, g->n[0], g->n[1], g->n[2], zthr, flags);
    return hip_check(hipGetLastError(), "indicator");
}

int acx_weno_judge_f32(const acx_grid *g, const float *HI, const uint8_t *used, const uint8_t *base, double thr,
                       double tol, uint8_t *out, long long *missed, double *worst, void *stream)
{
    int rc = check_ac4(g);
    if (rc) return rc;
    if (!HI || (!out && !missed) || (missed && !used)) return set_error(ACX_EINVAL, "judge: bad arguments");
    CwArgs p;
    memset(&p, 0, sizeof(p));
    p.nx = g->n[0]; p.ny = g->n[1]; p.nz = g->n[2];
    p.g4 = g->gamma_hi;
    p.g3c = (1.0 - g->gamma_hi) * g->gamma_lo;
    p.g3s = (1. / 6.) * (1 - g->gamma_hi) * (1 - g->gamma_lo);
    p.inv_g4 = 1.0 / g->gamma_hi;
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *d = nullptr;   // [0] misses, [1] low word: worst difference (float bits)
    if (missed) {
        if ((rc = hip_check(hipMallocAsync((void **)&d, 2 * sizeof(unsigned long long), s), "judge counter"))) return rc;
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, (float *)d, 4, 0.f);
    }
    const unsigned blocks = (unsigned)(((g->n[0] + JT_BX - 1) / JT_BX) * ((g->n[1] + JT_BY - 1) / JT_BY) *
                                       ((g->n[2] + JT_ZC - 1) / JT_ZC));
    hipLaunchKernelGGL(k_judge, dim3(blocks), dim3(JT_BX, JT_BY), 0, s, HI, p, (float)thr, (float)tol, used, base, out,
                       d, d ? (unsigned *)(d + 1) : nullptr);
    rc = hip_check(hipGetLastError(), "judge");
    if (missed) {
        unsigned long long h[2] = {0, 0};
        if (!rc) rc = hip_check(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s), "judge copy");
        const int rf = hip_check(hipFreeAsync(d, s), "judge counter free");
        if (!rc) rc = rf;
        if (!rc) rc = hip_check(hipStreamSynchronize(s), "judge sync");
        *missed = (long long)h[0];
        if (worst) {
            const unsigned bits = (unsigned)(h[1] & 0xffffffffu);
            float f;
            memcpy(&f, &bits, sizeof(f));
            *worst = f;
        }
    }
    return rc;
}

int acx_flags_dilate(const acx_grid *g, const uint8_t *in, uint8_t *out, void *stream)
{
    if (!g || !in || !out || in == out) return set_error(ACX_EINVAL, "dilate: bad arguments");
    const dim3 grd((unsigned)((g->n[0] + DIL_BX - 1) / DIL_BX), (unsigned)((g->n[1] + DIL_BY - 1) / DIL_BY),
                   (unsigned)((g->n[2] + DIL_ZC - 1) / DIL_ZC));
    hipLaunchKernelGGL(k_dilate, grd, dim3(DIL_BX, DIL_BY), 0, (hipStream_t)stream, in, out, g->n[0], g->n[1], g->n[2]);
    return hip_check(hipGetLastError(), "dilate");
}

int acx_flags_count(const uint8_t *flags, size_t ncells, long long *bits, void *stream)
{
    if (!flags || !bits) return set_error(ACX_EINVAL, "count: bad arguments");
    unsigned long long *d = nullptr;
    hipStream_t s = (hipStream_t)stream;
    int rc = hip_check(hipMallocAsync((void **)&d, sizeof(unsigned long long), s), "count alloc");
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, (float *)d, 2, 0.f);
    hipLaunchKernelGGL(k_count, dim3(1024), dim3(256), 0, s, flags, ncells, d);
    unsigned long long h = 0;
    rc = hip_check(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s), "count copy");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "count sync");
    (void)hipFreeAsync(d, s);
    *bits = (long long)h;
    return rc;
}

} // extern "C"
