/*
 * rhs3d.c -- oracle restatement of the 3D right-hand sides (test
 * infrastructure only; see oracle.h).
 *
 * Ghost handling follows the DMDA the reference builds
 * (3d/weno4/plane_wave/hype.c:87-104: BOX stencil, width 3, periodic axes
 * wrap, other axes DM_BOUNDARY_GHOSTED) and the adiabatic odd reflection of
 * component 0 applied pass by pass in x, y, z order
 * (3d/weno4/plane_wave/rhs_function.c:48-135).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define GW 3   /* ghost width, hype.c:104 */

/* 2x2 tensor Gauss rule on a face, hype.h:79-82 */
static const double gp_a[4] = {-0.28867513459481287, -0.28867513459481287, 0.28867513459481287, 0.28867513459481287};
static const double gp_b[4] = {-0.28867513459481287, 0.28867513459481287, -0.28867513459481287, 0.28867513459481287};
static const double gp_w[4] = {0.25, 0.25, 0.25, 0.25};

int orc_nvar(int system) { return system == ORC_WAVE2 ? 2 : 4; }

/* Local (ghosted) copy, (N+6)^3 cells, interlaced dof. */
static double *ghost_fill(const orc_grid3d *g, const double *U, int nv)
{
    const int Nx = g->N[0], Ny = g->N[1], Nz = g->N[2];
    const int Gx = Nx + 2 * GW, Gy = Ny + 2 * GW, Gz = Nz + 2 * GW;
    double *L = (double *)calloc((size_t)Gx * Gy * Gz * nv, sizeof(double));
    if (!L) return NULL;
    #pragma omp parallel for schedule(static)
    for (int kk = 0; kk < Gz; ++kk)
        for (int jj = 0; jj < Gy; ++jj)
            for (int ii = 0; ii < Gx; ++ii) {
                int i = ii - GW, j = jj - GW, k = kk - GW;
                int in_x = (i >= 0 && i < Nx), in_y = (j >= 0 && j < Ny), in_z = (k >= 0 && k < Nz);
                if (!in_x) { if (g->bc[0] != ORC_PERIODIC) continue; i = (i + Nx) % Nx; }
                if (!in_y) { if (g->bc[1] != ORC_PERIODIC) continue; j = (j + Ny) % Ny; }
                if (!in_z) { if (g->bc[2] != ORC_PERIODIC) continue; k = (k + Nz) % Nz; }
                const double *src = U + (((size_t)k * Ny + j) * Nx + i) * nv;
                double *dst = L + (((size_t)kk * Gy + jj) * Gx + ii) * nv;
                for (int c = 0; c < nv; ++c) dst[c] = src[c];
            }
    /* adiabatic walls: odd reflection of component 0, x then y then z
     * (rhs_function.c:48-135; irhs = -1-i on the low side, 2(N-1)-i+1 on the high side) */
#define AT(i, j, k) L[((((size_t)((k) + GW)) * Gy + ((j) + GW)) * Gx + ((i) + GW)) * nv]
    if (g->bc[0] == ORC_ADIABATIC)
        for (int k = -GW; k < Nz + GW; ++k)
            for (int j = -GW; j < Ny + GW; ++j)
                for (int i = -GW; i < Nx + GW; ++i) {
                    if (i < 0) AT(i, j, k) = -AT(-1 - i, j, k);
                    if (i >= Nx) AT(i, j, k) = -AT(2 * (Nx - 1) - i + 1, j, k);
                }
    if (g->bc[1] == ORC_ADIABATIC)
        for (int k = -GW; k < Nz + GW; ++k)
            for (int j = -GW; j < Ny + GW; ++j)
                for (int i = -GW; i < Nx + GW; ++i) {
                    if (j < 0) AT(i, j, k) = -AT(i, -1 - j, k);
                    if (j >= Ny) AT(i, j, k) = -AT(i, 2 * (Ny - 1) - j + 1, k);
                }
    if (g->bc[2] == ORC_ADIABATIC)
        for (int k = -GW; k < Nz + GW; ++k)
            for (int j = -GW; j < Ny + GW; ++j)
                for (int i = -GW; i < Nx + GW; ++i) {
                    if (k < 0) AT(i, j, k) = -AT(i, j, -1 - k);
                    if (k >= Nz) AT(i, j, k) = -AT(i, j, 2 * (Nz - 1) - k + 1);
                }
#undef AT
    /* LinearisedEuler walls (rhs_function.c:47-176): x pass, then y, then z,
     * each over the whole ghost box; all components copied from the mirror
     * cell (reflective: irhs = -1-i low, 2(N-1)-i+1 high, normal velocity
     * component 1+d negated) or from the boundary cell (transmissive). */
#define CELL(i, j, k) (L + ((((size_t)((k) + GW)) * Gy + ((j) + GW)) * Gx + ((i) + GW)) * nv)
    const int Nn[3] = {Nx, Ny, Nz};
    for (int d = 0; d < 3; ++d) {
        const int kind = g->bc[d];
        if (kind != ORC_REFLECTIVE && kind != ORC_TRANSMISSIVE) continue;
        for (int k = -GW; k < Nz + GW; ++k)
            for (int j = -GW; j < Ny + GW; ++j)
                for (int i = -GW; i < Nx + GW; ++i) {
                    const int q = d == 0 ? i : (d == 1 ? j : k), n = Nn[d];
                    if (q >= 0 && q < n) continue;
                    int r;
                    if (kind == ORC_TRANSMISSIVE) r = q < 0 ? 0 : n - 1;
                    else r = q < 0 ? -1 - q : 2 * (n - 1) - q + 1;
                    const int si = d == 0 ? r : i, sj = d == 1 ? r : j, sk = d == 2 ? r : k;
                    const double *src = CELL(si, sj, sk);
                    double *dst = CELL(i, j, k);
                    for (int c = 0; c < nv; ++c) dst[c] = src[c];
                    if (kind == ORC_REFLECTIVE && nv == 4) dst[1 + d] = -dst[1 + d];
                }
    }
#undef CELL
    return L;
}

/* face f of a cell: the Gauss point q in cell-local coordinates
 * (order of rhs_function.c:182-214: -x, +x, -y, +y, -z, +z) */
static void face_point(int f, int q, double *x, double *y, double *z)
{
    double a = gp_a[q], b = gp_b[q];
    switch (f) {
    case 0: *x = -0.5; *y = a; *z = b; break;
    case 1: *x = 0.5;  *y = a; *z = b; break;
    case 2: *x = a; *y = -0.5; *z = b; break;
    case 3: *x = a; *y = 0.5;  *z = b; break;
    case 4: *x = a; *y = b; *z = -0.5; break;
    default: *x = a; *y = b; *z = 0.5; break;
    }
}

/* PDEConsFlux + PDELLFRiemannSolver (3d/weno3/LinearisedEuler/pde.c:35-95);
 * the p<1e-12 abort (:50-55) is replaced by a NaN/Inf check by the caller. */
static double llf_acoustic(const double QL[4], const double QR[4], double nx, double ny, double nz,
                           double rho0, double c0, double flux[4])
{
    const double K = rho0 * c0 * c0, ir = 1.0 / rho0;
    double FL[4], FR[4];
    FL[0] = nx * K * QL[1] + ny * K * QL[2] + nz * K * QL[3];
    FL[1] = ir * nx * QL[0];
    FL[2] = ir * ny * QL[0];
    FL[3] = ir * nz * QL[0];
    FR[0] = nx * K * QR[1] + ny * K * QR[2] + nz * K * QR[3];
    FR[1] = ir * nx * QR[0];
    FR[2] = ir * ny * QR[0];
    FR[3] = ir * nz * QR[0];
    const double s = c0 > c0 ? c0 : c0;   /* max(s_L, s_R), pde.c:66,87 */
    for (int c = 0; c < 4; ++c) flux[c] = 0.5 * (FR[c] + FL[c] - s * (QR[c] - QL[c]));
    return s;
}

int orc_rhs3d_faithful(const orc_grid3d *g, const double *U, double *RHS, double *dt_out)
{
    const int nv = orc_nvar(g->system);
    const int Nx = g->N[0], Ny = g->N[1], Nz = g->N[2];
    const int Gx = Nx + 2 * GW, Gy = Ny + 2 * GW;
    const int Rx = Nx + 2, Ry = Ny + 2, Rz = Nz + 2;   /* reconstructed cells [-1, N] */
    const double r1_h = 1. / g->h;
    const size_t sx = nv, sy = (size_t)Gx * nv, sz = (size_t)Gx * Gy * nv;
    int bad = 0;

    double *L = ghost_fill(g, U, nv);
    if (!L) return 2;
    /* scratch: WAVE2 gradients [cell][face][gp][3] of comp 0 (hype.c:132),
     * ACOUSTIC4 values [cell][comp][face][gp] (LinearisedEuler/hype.c) */
    const size_t per_cell = (g->system == ORC_WAVE2) ? 6 * 4 * 3 : (size_t)nv * 6 * 4;
    double *S = (double *)malloc((size_t)Rx * Ry * Rz * per_cell * sizeof(double));
    double *Fx = (double *)malloc((size_t)Nz * Ny * (Nx + 1) * nv * sizeof(double));
    double *Fy = (double *)malloc((size_t)Nz * (Ny + 1) * Nx * nv * sizeof(double));
    double *Fz = (double *)malloc((size_t)(Nz + 1) * Ny * Nx * nv * sizeof(double));
    if (!S || !Fx || !Fy || !Fz) { free(L); free(S); free(Fx); free(Fy); free(Fz); return 2; }

#define LOC(i, j, k) (L + (((size_t)((k) + GW) * Gy + ((j) + GW)) * Gx + ((i) + GW)) * nv)
#define SCR(i, j, k) (S + (((size_t)((k) + 1) * Ry + ((j) + 1)) * Rx + ((i) + 1)) * per_cell)

    /* reconstruction loop (rhs_function.c:141-217 / LinearisedEuler :180-239) */
    #pragma omp parallel for schedule(static)
    for (int k = -1; k < Nz + 1; ++k)
        for (int j = -1; j < Ny + 1; ++j)
            for (int i = -1; i < Nx + 1; ++i) {
                double co[20];
                double *s = SCR(i, j, k);
                if (g->system == ORC_WAVE2) {
                    orc_cweno3d(LOC(i, j, k), sx, sy, sz, g->gamma_hi, g->gamma_lo, co, NULL);
                    for (int q = 0; q < 4; ++q)
                        for (int f = 0; f < 6; ++f) {
                            double x, y, z;
                            face_point(f, q, &x, &y, &z);
                            double *gr = s + (f * 4 + q) * 3;
                            orc_grad3d(co, x, y, z, g->h, gr, gr + 1, gr + 2);
                        }
                } else {
                    for (int c = 0; c < nv; ++c) {
                        orc_cweno3d(LOC(i, j, k) + c, sx, sy, sz, g->gamma_hi, g->gamma_lo, co, NULL);
                        for (int q = 0; q < 4; ++q)
                            for (int f = 0; f < 6; ++f) {
                                double x, y, z;
                                face_point(f, q, &x, &y, &z);
                                s[(c * 6 + f) * 4 + q] = orc_poly3d(x, y, z, co);
                            }
                    }
                }
            }

    /* flux sweeps (rhs_function.c:223-322 / LinearisedEuler :245-361) */
    double s_max = 0.0;
    for (int dir = 0; dir < 3; ++dir) {
        const int ex = (dir == 0), ey = (dir == 1), ez = (dir == 2);
        const double nx = ex, ny = ey, nz = ez;
        const int fl = 2 * dir + 1, fr = 2 * dir;   /* + face of left cell, - face of right cell */
        double *Fd = dir == 0 ? Fx : dir == 1 ? Fy : Fz;
        const int Mx = Nx + ex, My = Ny + ey, Mz = Nz + ez;
        #pragma omp parallel for schedule(static) reduction(max:s_max) reduction(|:bad)
        for (int k = 0; k < Mz; ++k)
            for (int j = 0; j < My; ++j)
                for (int i = 0; i < Mx; ++i) {
                    const double *sl = SCR(i - ex, j - ey, k - ez);
                    const double *sr = SCR(i, j, k);
                    double *out = Fd + (((size_t)k * My + j) * Mx + i) * nv;
                    if (g->system == ORC_WAVE2) {
                        double flux = 0.0;
                        for (int q = 0; q < 4; ++q) {
                            const double *gl = sl + (fl * 4 + q) * 3, *gr = sr + (fr * 4 + q) * 3;
                            /* PDEFlux + RiemannSolver, 3d/weno4/plane_wave/pde.c:12-34 */
                            double FL = g->c0 * g->c0 * (gl[0] * nx + gl[1] * ny + gl[2] * nz);
                            double FR = g->c0 * g->c0 * (gr[0] * nx + gr[1] * ny + gr[2] * nz);
                            flux += gp_w[q] * (0.5 * (FR + FL));
                        }
                        out[0] = flux;
                    } else {
                        double acc[4] = {0.0, 0.0, 0.0, 0.0};
                        for (int q = 0; q < 4; ++q) {
                            double QL[4], QR[4], F[4];
                            for (int c = 0; c < 4; ++c) {
                                QL[c] = sl[(c * 6 + fl) * 4 + q];
                                QR[c] = sr[(c * 6 + fr) * 4 + q];
                                if (!isfinite(QL[c]) || !isfinite(QR[c])) bad |= 1;
                            }
                            double s = llf_acoustic(QL, QR, nx, ny, nz, g->rho0, g->c0, F);
                            if (s > s_max) s_max = s;
                            for (int c = 0; c < 4; ++c) acc[c] += gp_w[q] * (F[c]);
                        }
                        for (int c = 0; c < 4; ++c) out[c] = acc[c];
                    }
                }
    }

    /* divergence (rhs_function.c:326-338 / LinearisedEuler :365-377) */
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                double *r = RHS + (((size_t)k * Ny + j) * Nx + i) * nv;
                const double *fx0 = Fx + (((size_t)k * Ny + j) * (Nx + 1) + i) * nv, *fx1 = fx0 + nv;
                const double *fy0 = Fy + (((size_t)k * (Ny + 1) + j) * Nx + i) * nv, *fy1 = fy0 + (size_t)Nx * nv;
                const double *fz0 = Fz + (((size_t)k * Ny + j) * Nx + i) * nv, *fz1 = fz0 + (size_t)Ny * Nx * nv;
                if (g->system == ORC_WAVE2) {
                    r[0] = LOC(i, j, k)[1];
                    r[1] = r1_h * (fx1[0] - fx0[0]) + r1_h * (fy1[0] - fy0[0]) + r1_h * (fz1[0] - fz0[0]);
                } else {
                    for (int c = 0; c < nv; ++c)
                        r[c] = -r1_h * (fx1[c] - fx0[c]) - r1_h * (fy1[c] - fy0[c]) - r1_h * (fz1[c] - fz0[c]);
                }
            }
#undef LOC
#undef SCR
    if (dt_out) *dt_out = (g->system == ORC_ACOUSTIC4 && s_max > 0.0) ? (g->cfl * g->h) / (3.0 * s_max) : 0.0;
    free(L); free(S); free(Fx); free(Fy); free(Fz);
    return bad;
}

/* ---------------------------------------------------------------------- */
/* Streaming form.  Face averages per cell and component:
 *   value  at x=+-1/2:  c0 +- c1/2 + c4/6 +- c10/20      (ACOUSTIC4)
 *   d/dx   at x=+-1/2: (c1 +- c4 + 3/5 c10)/h            (WAVE2)
 * and the analogues in y (c2,c5,c11) and z (c3,c6,c12).  Planes of face
 * averages are kept for z-1, z, z+1 while marching z. */

static void face_avgs(const orc_grid3d *g, const double *cell, size_t sx, size_t sy, size_t sz,
                      int nv, double *fa /* [nv][6] */)
{
    double co[20];
    if (g->system == ORC_WAVE2) {
        orc_cweno3d(cell, sx, sy, sz, g->gamma_hi, g->gamma_lo, co, NULL);
        const double ih = 1.0 / g->h;
        for (int d = 0; d < 3; ++d) {
            const double c1 = co[1 + d], c2 = co[4 + d], c3 = co[10 + d];
            fa[2 * d] = (c1 - c2 + 0.6 * c3) * ih;
            fa[2 * d + 1] = (c1 + c2 + 0.6 * c3) * ih;
        }
    } else {
        for (int c = 0; c < nv; ++c) {
            orc_cweno3d(cell + c, sx, sy, sz, g->gamma_hi, g->gamma_lo, co, NULL);
            for (int d = 0; d < 3; ++d) {
                const double c1 = co[1 + d], c2 = co[4 + d], c3 = co[10 + d];
                const double even = co[0] + c2 / 6.0, odd = 0.5 * c1 + c3 / 20.0;
                fa[c * 6 + 2 * d] = even - odd;
                fa[c * 6 + 2 * d + 1] = even + odd;
            }
        }
    }
}

static void face_flux(const orc_grid3d *g, int dir, const double *fl, const double *fr, int nv, double *F)
{
    /* fl: face averages of the left cell, fr: of the right cell ([nv][6]) */
    const int a = 2 * dir + 1, b = 2 * dir;
    if (g->system == ORC_WAVE2) {
        F[0] = g->c0 * g->c0 * 0.5 * (fl[a] + fr[b]);
    } else {
        double QL[4], QR[4];
        for (int c = 0; c < 4; ++c) { QL[c] = fl[c * 6 + a]; QR[c] = fr[c * 6 + b]; }
        llf_acoustic(QL, QR, dir == 0, dir == 1, dir == 2, g->rho0, g->c0, F);
    }
    (void)nv;
}

int orc_rhs3d_stream(const orc_grid3d *g, const double *U, double *RHS, int nthreads)
{
    const int nv = orc_nvar(g->system);
    const int Nx = g->N[0], Ny = g->N[1], Nz = g->N[2];
    const int Gx = Nx + 2 * GW, Gy = Ny + 2 * GW;
    const size_t sx = nv, sy = (size_t)Gx * nv, sz = (size_t)Gx * Gy * nv;
    const int Rx = Nx + 2, Ry = Ny + 2;
    const int nfa = (g->system == ORC_WAVE2) ? 6 : nv * 6;
    const size_t plane = (size_t)Rx * Ry * nfa;
    const double r1_h = 1.0 / g->h;
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double *L = ghost_fill(g, U, nv);
    if (!L) return 2;
#define LOC(i, j, k) (L + (((size_t)((k) + GW) * Gy + ((j) + GW)) * Gx + ((i) + GW)) * nv)
    #pragma omp parallel reduction(|:bad)
    {
        int nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads(); tid = omp_get_thread_num();
#endif
        const int k0 = (int)((long)Nz * tid / nt), k1 = (int)((long)Nz * (tid + 1) / nt);
        double *ring = (double *)malloc(3 * plane * sizeof(double));
        double Fa[4] = {0, 0, 0, 0}, Fb[4] = {0, 0, 0, 0};
        if (k1 > k0 && ring) {
            /* slot of plane k: (k+3) % 3 */
            for (int k = k0 - 1; k <= k1; ++k) {
                double *P = ring + (size_t)((k + 3) % 3) * plane;
                for (int j = -1; j <= Ny; ++j)
                    for (int i = -1; i <= Nx; ++i) {
                        /* only the face-neighbour halo is needed */
                        int halo = (i < 0 || i >= Nx) + (j < 0 || j >= Ny) + (k < k0 || k >= k1);
                        if (halo > 1) continue;
                        face_avgs(g, LOC(i, j, k), sx, sy, sz, nv, P + ((size_t)(j + 1) * Rx + (i + 1)) * nfa);
                    }
                if (k < k0 + 1) continue;
                const int kc = k - 1;   /* finish plane kc: needs kc-1, kc, kc+1 */
                const double *Pm = ring + (size_t)((kc + 2) % 3) * plane;
                const double *Pc = ring + (size_t)((kc + 3) % 3) * plane;
                const double *Pp = ring + (size_t)((kc + 4) % 3) * plane;
#define FA(P, i, j) ((P) + ((size_t)((j) + 1) * Rx + ((i) + 1)) * nfa)
                for (int j = 0; j < Ny; ++j)
                    for (int i = 0; i < Nx; ++i) {
                        double acc[4] = {0, 0, 0, 0};
                        const double *me = FA(Pc, i, j);
                        /* x */
                        face_flux(g, 0, FA(Pc, i - 1, j), me, nv, Fa);
                        face_flux(g, 0, me, FA(Pc, i + 1, j), nv, Fb);
                        for (int c = 0; c < 4; ++c) acc[c] += Fb[c] - Fa[c];
                        face_flux(g, 1, FA(Pc, i, j - 1), me, nv, Fa);
                        face_flux(g, 1, me, FA(Pc, i, j + 1), nv, Fb);
                        for (int c = 0; c < 4; ++c) acc[c] += Fb[c] - Fa[c];
                        face_flux(g, 2, FA(Pm, i, j), me, nv, Fa);
                        face_flux(g, 2, me, FA(Pp, i, j), nv, Fb);
                        for (int c = 0; c < 4; ++c) acc[c] += Fb[c] - Fa[c];
                        double *r = RHS + (((size_t)kc * Ny + j) * Nx + i) * nv;
                        if (g->system == ORC_WAVE2) {
                            r[0] = LOC(i, j, kc)[1];
                            r[1] = r1_h * acc[0];
                        } else {
                            for (int c = 0; c < nv; ++c) {
                                r[c] = -r1_h * acc[c];
                                if (!isfinite(r[c])) bad |= 1;
                            }
                        }
                    }
#undef FA
            }
        }
        free(ring);
    }
#undef LOC
    free(L);
    return bad;
}

int orc_rhs3d_cb_faithful(void *ctx, double t, const double *U, double *F)
{
    (void)t;
    return orc_rhs3d_faithful((const orc_grid3d *)ctx, U, F, NULL);
}

double orc_weno_source(double t, double f0)
{
    /* 3d/weno4/point_source/source.c:11-17 */
    double t0 = 4.0 / f0;
    return -2.0 * (t - t0) * f0 * f0 * exp(-1.0 * f0 * f0 * (t - t0) * (t - t0));
}

int orc_rhs3d_cb_point_source(void *ctx, double t, const double *U, double *F)
{
    orc_point_source3d *ps = (orc_point_source3d *)ctx;
    const orc_grid3d *g = &ps->g;
    const int nv = orc_nvar(g->system);
    int rc = orc_rhs3d_faithful(g, U, F, NULL);
    if (rc) return rc;
    const size_t nx = (size_t)g->N[0], ny = (size_t)g->N[1];
    /* rhs_function.c:337-340: rhs[isz][isy][isx].comp[1] += Source(t) * r1_h3 */
    const size_t s = (((size_t)ps->src[2] * ny + ps->src[1]) * nx + ps->src[0]) * nv + 1;
    F[s] += orc_weno_source(t, ps->f0) / (g->h * g->h * g->h);
    /* rhs_function.c:342-346: receiver u[irz][iry][irx].comp[0] at every RHS call */
    if (ps->ntrace < ps->cap) {
        const size_t r = (((size_t)ps->rcv[2] * ny + ps->rcv[1]) * nx + ps->rcv[0]) * nv;
        ps->trace_t[ps->ntrace] = t;
        ps->trace_u[ps->ntrace] = U[r];
        ps->ntrace++;
    }
    return 0;
}

int orc_rhs3d_cb_stream(void *ctx, double t, const double *U, double *F)
{
    (void)t;
    return orc_rhs3d_stream((const orc_grid3d *)ctx, U, F, 0);
}
