/*
 * CPU restatement of ICAR's MPDATA advection step (adv=2).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
 * checked against.  Only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark load it; the product library never links it and has no
 * CPU fallback.
 *
 * It restates, in plain C with the same operation order, the Fortran in
 *   src/physics/adv_mpdata.f90            (flux1, upwind_advection,
 *                                          mpdata_fluxes, flux_limiter,
 *                                          mpdata_advect3d, test_divergence,
 *                                          mpdata_compute_wind)
 *   src/physics/adv_mpdata_FCT_core.f90   (the 1-D FCT core)
 *   src/physics/advection_driver.f90:59-126 (advect: wind once, then species)
 * Every Fortran expression is evaluated left to right exactly as written, with
 * no contraction (build with -ffp-contract=off), so that the FP32 build is the
 * bit pattern a strict IEEE compile of the reference produces.  The same file
 * compiled with -DORACLE_DOUBLE=1 restates a build of the reference with
 * 8-byte default reals (literal constants promoted as well).
 *
 * Layout (shared with the HIP library): every 3-D field is Fortran
 * column-major (ims:ime, kms:kme, jms:jme) -> 0-based (i, k, j), i fastest:
 *     IX(i,k,j) = i + nx*(k + nz*j)
 * Face fields use the same nx*nz*ny shape:
 *     U[IX(i,k,j)]  x face between cells i-1 and i   (i >= 1; slot i=0 unused)
 *     V[IX(i,k,j)]  y face between rows  j-1 and j   (j >= 1; slot j=0 unused)
 *     W[IX(i,k,j)]  z face on top of level k          (k = nz-1 is the model top)
 * The domain's staggered inputs keep ICAR's shapes: u and jacobian_u are
 * (nx+1, nz, ny), v and jacobian_v are (nx, nz, ny+1).
 *
 * Parity status of each function is recorded in the test-suite headers
 * (tests/test_oracle_*.py): every function below is pinned by closed forms,
 * invariants or special cases there.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#if ORACLE_DOUBLE
typedef double real;
#define L(x) (x)
#else
typedef float real;
#define L(x) (x##f)
#endif

#define IX(i, k, j) ((size_t)(i) + (size_t)nx * ((size_t)(k) + (size_t)nz * (size_t)(j)))

static inline real rmax2(real a, real b) { return a > b ? a : b; }
static inline real rmin2(real a, real b) { return a < b ? a : b; }

/* Fortran ABS: clear the sign bit (|-0| = +0). */
static inline real fabs_r(real x)
{
#if ORACLE_DOUBLE
    return fabs(x);
#else
    return fabsf(x);
#endif
}

/* flux1, adv_mpdata.f90:23-41:  f = ((U+|U|)*l + (U-|U|)*r)/2 */
static inline real flux1(real l, real r, real U)
{
    return ((U + fabs_r(U)) * l + (U - fabs_r(U)) * r) / L(2.0);
}

/* ------------------------------------------------------------------------- */
/* mpdata_compute_wind, adv_mpdata.f90:477-504                               */
/* ------------------------------------------------------------------------- */
void oracle_compute_wind(int nx, int ny, int nz, real dx, real dt, int advect_density,
                         const real *u, const real *v, const real *w, const real *density,
                         const real *jaco_u, const real *jaco_v, const real *jaco_w,
                         real *U, real *V, real *W)
{
    const size_t N = (size_t)nx * nz * ny;
    real *rho = (real *)malloc(N * sizeof(real));
    /* :489-490  rho = 1, or the density when advect_density */
    for (size_t n = 0; n < N; n++) rho[n] = advect_density ? density[n] : L(1.0);

    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            /* :492-493  U_m(ims+1:ime) = u*dt*(rho_r+rho_l)*0.5*jacobian_u/dx */
            U[IX(0, k, j)] = 0;
            for (int i = 1; i < nx; i++) {
                size_t iu = (size_t)i + (size_t)(nx + 1) * ((size_t)k + (size_t)nz * j);
                U[IX(i, k, j)] = u[iu] * dt * (rho[IX(i, k, j)] + rho[IX(i - 1, k, j)]) * L(0.5) *
                                 jaco_u[iu] / dx;
            }
            /* :494-495  V_m(:,:,jms+1:jme), same form along y */
            for (int i = 0; i < nx; i++) {
                if (j == 0) {
                    V[IX(i, k, 0)] = 0;
                    continue;
                }
                size_t iv = IX(i, k, j); /* v has ny+1 rows, same stride for j <= ny */
                V[IX(i, k, j)] = v[iv] * dt * (rho[IX(i, k, j)] + rho[IX(i, k, j - 1)]) * L(0.5) *
                                 jaco_v[iv] / dx;
            }
            /* :496-498  W_m = w*dt*jacobian_w*(rho_up+rho)*0.5 ; top: w*dt*jacobian_w*rho */
            for (int i = 0; i < nx; i++) {
                size_t c = IX(i, k, j);
                if (k < nz - 1)
                    W[c] = w[c] * dt * jaco_w[c] * (rho[IX(i, k + 1, j)] + rho[c]) * L(0.5);
                else
                    W[c] = w[c] * dt * jaco_w[c] * rho[c];
            }
        }
    free(rho);
}

/* ------------------------------------------------------------------------- */
/* upwind_advection, adv_mpdata.f90:43-103 (donor cell, ring copied)        */
/* ------------------------------------------------------------------------- */
void oracle_upwind(int nx, int ny, int nz, const real *qin, const real *U, const real *V,
                   const real *W, const real *rho, real *q, const real *dz, const real *jaco)
{
    const size_t N = (size_t)nx * nz * ny;
    /* :61-63  q = qin on every row (the outer ring is never updated) */
    memcpy(q, qin, N * sizeof(real));
    for (int j = 1; j < ny - 1; j++) {
        for (int k = 0; k < nz; k++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IX(i, k, j);
                /* :72-79  x faces (f1), north face (f3), south face (f4) */
                real f1r = flux1(qin[c], qin[IX(i + 1, k, j)], U[IX(i + 1, k, j)]);
                real f1l = flux1(qin[IX(i - 1, k, j)], qin[c], U[c]);
                real f3 = flux1(qin[c], qin[IX(i, k, j + 1)], V[IX(i, k, j + 1)]);
                real f4 = flux1(qin[IX(i, k, j - 1)], qin[c], V[c]);
                /* :86-87  horizontal update divided by jaco*rho */
                q[c] = q[c] - ((f1r - f1l) + (f3 - f4)) / (jaco[c] * rho[c]);
            }
        for (int k = 0; k < nz; k++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IX(i, k, j);
                real den = dz[c] * jaco[c] * rho[c];
                if (k == 0) {
                    /* :93-94  no flux through the ground */
                    real f5 = flux1(qin[c], qin[IX(i, 1, j)], W[c]);
                    q[c] = q[c] - f5 / den;
                } else if (k == nz - 1) {
                    /* :96-97  top flux qin(kme)*w(kme), not upwinded */
                    real f5m = flux1(qin[IX(i, k - 1, j)], qin[c], W[IX(i, k - 1, j)]);
                    q[c] = q[c] - (qin[c] * W[c] - f5m) / den;
                } else {
                    /* :90-91  interior levels */
                    real f5 = flux1(qin[c], qin[IX(i, k + 1, j)], W[c]);
                    real f5m = flux1(qin[IX(i, k - 1, j)], qin[c], W[IX(i, k - 1, j)]);
                    q[c] = q[c] - (f5 - f5m) / den;
                }
            }
    }
}

/* ------------------------------------------------------------------------- */
/* mpdata_fluxes, adv_mpdata.f90:105-252 (antidiffusive pseudo-velocities)   */
/*  q: field after the previous pass; U,V: Courant numbers; Wd = W_m/dz;     */
/*  G = jaco*rho.                                                            */
/* ------------------------------------------------------------------------- */
void oracle_fluxes(int nx, int ny, int nz, const real *q, const real *U, const real *V,
                   const real *Wd, real *u2, real *v2, real *w2, const real *G)
{
    const size_t N = (size_t)nx * nz * ny;
    const real eps = L(1e-10);
    /* :119-121 */
    memset(u2, 0, N * sizeof(real));
    memset(v2, 0, N * sizeof(real));
    memset(w2, 0, N * sizeof(real));
#define Q(a, b, c) q[IX(a, b, c)]
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            /* ---- U component, :131-165 (the guard `i > 0` is always true) */
            for (int i = 1; i < nx; i++) {
                size_t c = IX(i, k, j), cl = IX(i - 1, k, j);
                real rx = q[c], lx = q[cl];
                real denom = rx + lx + eps;
                real au = fabs_r(U[c]);
                real val = au * (1 - au / (L(0.5) * (G[c] + G[cl])));
                val = val * (rx - lx) / denom;
                if (j > 0 && j < ny - 1) { /* UxV, :145-152 */
                    real eq = (Q(i, k, j + 1) - Q(i, k, j - 1) + Q(i - 1, k, j + 1) - Q(i - 1, k, j - 1)) /
                              (Q(i, k, j + 1) + Q(i, k, j - 1) + Q(i - 1, k, j + 1) + Q(i - 1, k, j - 1) + eps);
                    real ev = (L(1.0) / L(4.0)) *
                              (V[IX(i, k, j)] + V[IX(i, k, j + 1)] + V[IX(i - 1, k, j)] + V[IX(i - 1, k, j + 1)]);
                    val = val - L(0.5) * U[c] * ev * eq / (G[c] + G[cl]);
                }
                if (k > 0 && k < nz - 1) { /* UxW, :157-164 */
                    real eq = (Q(i, k + 1, j) - Q(i, k - 1, j) + Q(i - 1, k + 1, j) - Q(i - 1, k - 1, j)) /
                              (Q(i, k + 1, j) + Q(i, k - 1, j) + Q(i - 1, k + 1, j) + Q(i - 1, k - 1, j) + eps);
                    real ev = (L(1.0) / L(4.0)) *
                              (Wd[IX(i, k, j)] + Wd[IX(i, k - 1, j)] + Wd[IX(i - 1, k, j)] + Wd[IX(i - 1, k - 1, j)]);
                    val = val - L(0.5) * U[c] * ev * eq / (G[c] + G[cl]);
                }
                u2[c] = val;
            }
            /* ---- V component, :169-205 */
            if (j > 0)
                for (int i = 0; i < nx; i++) {
                    size_t c = IX(i, k, j), cs = IX(i, k, j - 1);
                    real r = q[c], l = q[cs];
                    real denom = r + l + eps;
                    real av = fabs_r(V[c]);
                    real val = av * (1 - av / (L(0.5) * (G[c] + G[cs])));
                    val = val * (r - l) / denom;
                    if (i > 0 && i < nx - 1) { /* VxU, :186-192 (edge_v/edge_q zero at the ends) */
                        real eq = (Q(i + 1, k, j - 1) - Q(i - 1, k, j) + Q(i + 1, k, j) - Q(i - 1, k, j - 1)) /
                                  (Q(i + 1, k, j) + Q(i + 1, k, j - 1) + Q(i - 1, k, j) + Q(i - 1, k, j - 1) + eps);
                        real ev = (L(1.0) / L(4.0)) *
                                  (U[IX(i + 1, k, j)] + U[IX(i + 1, k, j - 1)] + U[IX(i, k, j)] + U[IX(i, k, j - 1)]);
                        val = val - L(0.5) * V[c] * ev * eq / (G[c] + G[cs]);
                    }
                    if (k > 0 && k < nz - 1) { /* VxW, :197-204 */
                        real eq = (Q(i, k + 1, j - 1) - Q(i, k - 1, j) + Q(i, k + 1, j) - Q(i, k - 1, j - 1)) /
                                  (Q(i, k + 1, j - 1) + Q(i, k - 1, j) + Q(i, k + 1, j) + Q(i, k - 1, j - 1) + eps);
                        real ev = (L(1.0) / L(4.0)) *
                                  (Wd[IX(i, k, j)] + Wd[IX(i, k - 1, j)] + Wd[IX(i, k, j - 1)] + Wd[IX(i, k - 1, j - 1)]);
                        val = val - L(0.5) * V[c] * ev * eq / (G[c] + G[cs]);
                    }
                    v2[c] = val;
                }
            /* ---- W component, :211-246 (w2 = 0 at the top level) */
            for (int i = 0; i < nx; i++) {
                size_t c = IX(i, k, j);
                if (k == nz - 1) {
                    w2[c] = 0;
                    continue;
                }
                size_t cu = IX(i, k + 1, j);
                real r = q[cu], l = q[c];
                real denom = r + l + eps;
                real aw = fabs_r(Wd[c]);
                real val = aw * (1 - aw / (L(0.5) * (G[cu] + G[c])));
                val = val * (r - l) / denom;
                if (i > 0 && i < nx - 1) { /* WxU, :227-233 */
                    real eq = (Q(i + 1, k + 1, j) - Q(i - 1, k, j) + Q(i + 1, k, j) - Q(i - 1, k + 1, j)) /
                              (Q(i + 1, k, j) + Q(i + 1, k + 1, j) + Q(i - 1, k, j) + Q(i - 1, k + 1, j) + eps);
                    real ev = (L(1.0) / L(4.0)) *
                              (U[IX(i + 1, k, j)] + U[IX(i + 1, k + 1, j)] + U[IX(i, k, j)] + U[IX(i, k + 1, j)]);
                    val = val - L(0.5) * Wd[c] * ev * eq / (G[cu] + G[c]);
                }
                if (j > 0 && j < ny - 1) { /* WxV, :238-245 */
                    real eq = (Q(i, k + 1, j + 1) - Q(i, k, j - 1) + Q(i, k, j + 1) - Q(i, k + 1, j - 1)) /
                              (Q(i, k, j + 1) + Q(i, k + 1, j - 1) + Q(i, k + 1, j + 1) + Q(i, k, j - 1) + eps);
                    real ev = (L(1.0) / L(4.0)) *
                              (V[IX(i, k, j)] + V[IX(i, k + 1, j)] + V[IX(i, k, j + 1)] + V[IX(i, k + 1, j + 1)]);
                    val = val - L(0.5) * Wd[c] * ev * eq / (G[cu] + G[c]);
                }
                w2[c] = val;
            }
        }
#undef Q
}

/* ------------------------------------------------------------------------- */
/* The 1-D FCT core, adv_mpdata_FCT_core.f90:47-116.                         */
/*   n cells (0..n-1), n-1 faces; U2[t] is the face between cells t and t+1. */
/*   q1 = field after the previous pass, l = field before it (q0).           */
/*   f is scratch of n-1 entries.  Requires n >= 3.                          */
/* ------------------------------------------------------------------------- */
void oracle_fct_core(int n, const real *q1, const real *l, real *U2, int flux_is_w, real *f)
{
    const real eps = L(1e-15);
    /* flux1(q1(n_s:n-1), q1(n_s+1:n), U2, f), adv_mpdata.f90:296,314,341 */
    for (int t = 0; t < n - 1; t++) f[t] = flux1(q1[t], q1[t + 1], U2[t]);
    real qmax_i = 0, qmin_i = 0, qmax_i2 = 0, qmin_i2 = 0;
    real fin_i = 0, fout_i = 0, fin_i2 = 0, fout_i2 = 0;
    const int last = n - 2;
    for (int i = 0; i <= last; i++) {
        if (i == 0) { /* :50-55 */
            qmax_i = rmax2(rmax2(q1[i], q1[i + 1]), rmax2(l[i], l[i + 1]));
            qmin_i = rmin2(rmin2(q1[i], q1[i + 1]), rmin2(l[i], l[i + 1]));
            qmax_i2 = rmax2(rmax2(rmax2(q1[i], q1[i + 1]), rmax2(q1[i + 2], l[i])), rmax2(l[i + 1], l[i + 2]));
            qmin_i2 = rmin2(rmin2(rmin2(q1[i], q1[i + 1]), rmin2(q1[i + 2], l[i])), rmin2(l[i + 1], l[i + 2]));
        } else if (i != last) { /* :56-61 */
            qmax_i = qmax_i2;
            qmin_i = qmin_i2;
            qmax_i2 = rmax2(rmax2(rmax2(q1[i], q1[i + 1]), rmax2(q1[i + 2], l[i])), rmax2(l[i + 1], l[i + 2]));
            qmin_i2 = rmin2(rmin2(rmin2(q1[i], q1[i + 1]), rmin2(q1[i + 2], l[i])), rmin2(l[i + 1], l[i + 2]));
        } else { /* :62-68  the last cell's window leaves out l(n) */
            qmax_i = qmax_i2;
            qmin_i = qmin_i2;
            qmax_i2 = rmax2(rmax2(q1[i], q1[i + 1]), l[i]);
            qmin_i2 = rmin2(rmin2(q1[i], q1[i + 1]), l[i]);
        }
        /* :72-85 */
        if (i != 0) {
            fin_i = fin_i2;
            fout_i = fout_i2;
        } else if (flux_is_w) {
            fin_i = L(0.) - rmin2(L(0.), f[i]);
            fout_i = rmax2(L(0.), f[i]);
        } else {
            fin_i = 0;
            fout_i = 0;
        }
        /* :88-100 */
        if (i != last) {
            fin_i2 = rmax2(L(0.), f[i]) - rmin2(L(0.), f[i + 1]);
            fout_i2 = rmax2(L(0.), f[i + 1]) - rmin2(L(0.), f[i]);
        } else if (flux_is_w) {
            fin_i2 = rmax2(L(0.), f[i]) - rmin2(L(0.), f[i]);
            fout_i2 = rmax2(L(0.), f[i]) - rmin2(L(0.), f[i]);
        } else {
            fin_i2 = 0;
            fout_i2 = 0;
        }
        /* :103-115 */
        if (U2[i] > 0) {
            real beta_out_i = (q1[i] - qmin_i) / (fout_i + eps);
            real beta_in_i2 = (qmax_i2 - q1[i + 1]) / (fin_i2 + eps);
            U2[i] = rmin2(rmin2(L(1.), beta_in_i2), beta_out_i) * U2[i];
        } else if (U2[i] < 0) {
            real beta_in_i = (qmax_i - q1[i]) / (fin_i + eps);
            real beta_out_i2 = (q1[i + 1] - qmin_i2) / (fout_i2 + eps);
            U2[i] = rmin2(rmin2(L(1.), beta_in_i), beta_out_i2) * U2[i];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* flux_limiter, adv_mpdata.f90:254-350                                       */
/*   q0 = field before pass 1 (dummy q), q1 = field after it (dummy q2)      */
/* ------------------------------------------------------------------------- */
void oracle_flux_limiter(int nx, int ny, int nz, const real *q0, const real *q1, real *u2, real *v2,
                         real *w2)
{
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    real *a = (real *)malloc(sizeof(real) * nmax);
    real *b = (real *)malloc(sizeof(real) * nmax);
    real *U2 = (real *)malloc(sizeof(real) * nmax);
    real *f = (real *)malloc(sizeof(real) * nmax);
    for (int j = 1; j < ny - 1; j++) {
        /* :283-300  x rows, all levels; dummy u(ims:ime-1) is u2(ims+1:ime) */
        for (int k = 0; k < nz; k++) {
            for (int i = 0; i < nx; i++) {
                a[i] = q1[IX(i, k, j)];
                b[i] = q0[IX(i, k, j)];
            }
            for (int t = 0; t < nx - 1; t++) U2[t] = u2[IX(t + 1, k, j)];
            oracle_fct_core(nx, a, b, U2, 0, f);
            for (int t = 0; t < nx - 1; t++) u2[IX(t + 1, k, j)] = U2[t];
        }
        /* :302-319  z columns of the interior x range, then w(kme) = 0 */
        for (int i = 1; i < nx - 1; i++) {
            for (int k = 0; k < nz; k++) {
                a[k] = q1[IX(i, k, j)];
                b[k] = q0[IX(i, k, j)];
            }
            for (int t = 0; t < nz - 1; t++) U2[t] = w2[IX(i, t, j)];
            oracle_fct_core(nz, a, b, U2, 1, f);
            for (int t = 0; t < nz - 1; t++) w2[IX(i, t, j)] = U2[t];
            w2[IX(i, nz - 1, j)] = 0;
        }
    }
    /* :324-346  y rows for every x and level; dummy v(jms:jme-1) is v2(jms+1:jme) */
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            for (int j = 0; j < ny; j++) {
                a[j] = q1[IX(i, k, j)];
                b[j] = q0[IX(i, k, j)];
            }
            for (int t = 0; t < ny - 1; t++) U2[t] = v2[IX(i, k, t + 1)];
            oracle_fct_core(ny, a, b, U2, 0, f);
            for (int t = 0; t < ny - 1; t++) v2[IX(i, k, t + 1)] = U2[t];
        }
    free(a);
    free(b);
    free(U2);
    free(f);
}

/* ------------------------------------------------------------------------- */
/* mpdata_advect3d, adv_mpdata.f90:352-413                                    */
/*   rho is the dummy argument (ICAR passes domain%density even when         */
/*   advect_density is off); U,V,W are the Courant numbers of compute_wind.  */
/* ------------------------------------------------------------------------- */
int oracle_advect3d(int nx, int ny, int nz, real *q, const real *rho, const real *jaco, const real *dz,
                    const real *U, const real *V, const real *W, int order, int fct)
{
    if (nx < 3 || ny < 3 || nz < 3) return -1;
    const size_t N = (size_t)nx * nz * ny;
    real *q2 = (real *)malloc(N * sizeof(real));
    real *u2 = (real *)malloc(N * sizeof(real));
    real *v2 = (real *)malloc(N * sizeof(real));
    real *w2 = (real *)malloc(N * sizeof(real));
    real *Wd = (real *)malloc(N * sizeof(real));
    real *G = (real *)malloc(N * sizeof(real));
    if (!q2 || !u2 || !v2 || !w2 || !Wd || !G) {
        free(q2); free(u2); free(v2); free(w2); free(Wd); free(G);
        return -2;
    }
    for (int iord = 1; iord <= order; iord++) {
        if (iord == 1) {
            oracle_upwind(nx, ny, nz, q, U, V, W, rho, q2, dz, jaco); /* :369 */
        } else {
            /* :374  temporaries W_m/dz and jaco*rho */
            for (size_t n = 0; n < N; n++) {
                Wd[n] = W[n] / dz[n];
                G[n] = jaco[n] * rho[n];
            }
            oracle_fluxes(nx, ny, nz, q2, U, V, Wd, u2, v2, w2, G);
            /* :378-380  damping by 0.5; w re-scaled by dz */
            for (size_t n = 0; n < N; n++) {
                u2[n] = u2[n] * L(0.5);
                v2[n] = v2[n] * L(0.5);
                w2[n] = w2[n] * L(0.5) * dz[n];
            }
            if (fct) oracle_flux_limiter(nx, ny, nz, q, q2, u2, v2, w2); /* :381-383 */
            oracle_upwind(nx, ny, nz, q2, u2, v2, w2, rho, q, dz, jaco); /* :384 */
        }
        if (iord != order) {
            if (iord > 1) memcpy(q2, q, N * sizeof(real)); /* :389-397 */
        } else if (iord == 1) {
            memcpy(q, q2, N * sizeof(real)); /* :399-407 */
        }
    }
    free(q2); free(u2); free(v2); free(w2); free(Wd); free(G);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* advect, advection_driver.f90:59-126 (MPDATA branch)                        */
/*   species: array of nspecies pointers, advanced in order, in place.       */
/*   With rk3 set MPDATA does nothing (:112-115).                            */
/* ------------------------------------------------------------------------- */
int oracle_advect(int nx, int ny, int nz, real dx, real dt, int order, int fct, int advect_density,
                  int rk3, const real *u, const real *v, const real *w, const real *density,
                  const real *jaco, const real *jaco_u, const real *jaco_v, const real *jaco_w,
                  const real *dz, int nspecies, real **species)
{
    if (nx < 3 || ny < 3 || nz < 3) return -1;
    const size_t N = (size_t)nx * nz * ny;
    real *U = (real *)malloc(N * sizeof(real));
    real *V = (real *)malloc(N * sizeof(real));
    real *W = (real *)malloc(N * sizeof(real));
    if (!U || !V || !W) {
        free(U); free(V); free(W);
        return -2;
    }
    oracle_compute_wind(nx, ny, nz, dx, dt, advect_density, u, v, w, density, jaco_u, jaco_v, jaco_w,
                        U, V, W);
    int rc = 0;
    if (!rk3)
        for (int s = 0; s < nspecies && rc == 0; s++)
            rc = oracle_advect3d(nx, ny, nz, species[s], density, jaco, dz, U, V, W, order, fct);
    free(U); free(V); free(W);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* test_divergence, adv_mpdata.f90:445-475: number of interior cells whose   */
/* discrete divergence of the Courant numbers exceeds 1e-3, and the maximum. */
/* ------------------------------------------------------------------------- */
long oracle_divergence(int nx, int ny, int nz, const real *U, const real *V, const real *W,
                       const real *dz, double *max_abs)
{
    long count = 0;
    double mx = 0;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 0; k < nz; k++) {
                real du = U[IX(i + 1, k, j)] - U[IX(i, k, j)];
                real dv = V[IX(i, k, j + 1)] - V[IX(i, k, j)];
                real dw = (k == 0) ? W[IX(i, k, j)] / dz[IX(i, k, j)]
                                   : (W[IX(i, k, j)] - W[IX(i, k - 1, j)]) / dz[IX(i, k, j)];
                real d = fabs_r(du + dv + dw);
                if (d > L(1e-3)) count++;
                if ((double)d > mx) mx = (double)d;
            }
    if (max_abs) *max_abs = mx;
    return count;
}

/* Mass weight sum(q * jaco * rho * dz) over the interior (the cells advect  */
/* updates), accumulated in double.  In the flux form of upwind_advection    */
/* (adv_mpdata.f90:86-97) with per-level dz this total changes only through  */
/* the fluxes across the faces of the outer ring.                            */
double oracle_mass(int nx, int ny, int nz, const real *q, const real *jaco, const real *rho,
                   const real *dz)
{
    double s = 0;
    for (int j = 1; j < ny - 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IX(i, k, j);
                s += (double)q[c] * (double)jaco[c] * (double)rho[c] * (double)dz[c];
            }
    return s;
}

/* Lateral boundary conditions applied to the outer ring between advect      */
/* calls (ICAR fills the ring through forcing and halo exchange; the tests    */
/* fill it the way test_mpdata.f90:96-97 does).  bc_x: 0 fixed ring,          */
/* 1 zero-gradient outflow at i = nx-1, 2 periodic.  bc_y: 0 fixed, 2         */
/* periodic (row 0 <- row ny-2, row ny-1 <- row 1).  x is applied first.      */
void oracle_apply_bc(int nx, int ny, int nz, real *q, int bc_x, int bc_y)
{
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            if (bc_x == 1) q[IX(nx - 1, k, j)] = q[IX(nx - 2, k, j)];
            if (bc_x == 2) {
                q[IX(0, k, j)] = q[IX(nx - 2, k, j)];
                q[IX(nx - 1, k, j)] = q[IX(1, k, j)];
            }
        }
    if (bc_y == 2)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                q[IX(i, k, 0)] = q[IX(i, k, ny - 2)];
                q[IX(i, k, ny - 1)] = q[IX(i, k, 1)];
            }
}

int oracle_real_size(void) { return (int)sizeof(real); }
