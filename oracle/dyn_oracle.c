/*
 * CPU restatement of the steps ICAR runs next to advect: the wind balance,
 * the time-step bound and the lateral forcing of advected species.
 *
 * TEST INFRASTRUCTURE ONLY, like mpdata_oracle.c (same build, same real
 * type): the HIP library is checked against it and never links it.
 *
 *   balance_uvw (update=.false.), src/physics/wind.f90:88-164, as called
 *   before advection when advect_density is set (src/main/time_step.f90:324):
 *     calc_divergence(horz_only=.true.)   wind.f90:201-246 (horizontal part)
 *     calc_w                               wind.f90:166-199
 *   compute_dt, src/main/time_step.f90:51-168, with use_density=.false. as
 *   update_dt passes it (:229-242): its density branches are commented out
 *   in the reference and are not restated.
 *   init_relax_filters, src/objects/domain_obj.f90:3285-3390, and the 3-D
 *   boundary-forced branch of apply_forcing, :3493-3503 (called at
 *   src/main/time_step.f90:319 before the physics and advect).
 *
 * Layout as mpdata_oracle.c: IX(i,k,j) = i + nx*(k + nz*j); u and jacobian_u
 * are (nx+1, nz, ny), v and jacobian_v (nx, nz, ny+1).  Every expression is
 * evaluated left to right as the Fortran writes it (-ffp-contract=off).
 */
#include <math.h>
#include <stddef.h>

#if ORACLE_DOUBLE
typedef double real;
#define L(x) (x)
#define SQRT sqrt
#else
typedef float real;
#define L(x) (x##f)
#define SQRT sqrtf
#endif

#define IX(i, k, j) ((size_t)(i) + (size_t)nx * ((size_t)(k) + (size_t)nz * (size_t)(j)))
#define IXU(i, k, j) ((size_t)(i) + (size_t)(nx + 1) * ((size_t)(k) + (size_t)nz * (size_t)(j)))

static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real fabs_r(real x)
{
#if ORACLE_DOUBLE
    return fabs(x);
#else
    return fabsf(x);
#endif
}

/* u_met of x face i (0..nx) of row (k, j), wind.f90:222-228 */
static real u_met(int nx, int ny, int nz, int i, int k, int j, int adv_den, const real *u,
                  const real *ju, const real *rho)
{
    (void)ny;
    const size_t f = IXU(i, k, j);
    if (!adv_den) return u[f] * ju[f];
    if (i == 0) return u[f] * ju[f] * (L(1.5) * rho[IX(0, k, j)] - L(0.5) * rho[IX(1, k, j)]);
    if (i == nx)
        return u[f] * ju[f] * (L(1.5) * rho[IX(nx - 1, k, j)] - L(0.5) * rho[IX(nx - 2, k, j)]);
    return u[f] * ju[f] * (rho[IX(i - 1, k, j)] + rho[IX(i, k, j)]) / L(2.0);
}

/* v_met of y face j (0..ny) of column (i, k), wind.f90:225-227 */
static real v_met(int nx, int ny, int nz, int i, int k, int j, int adv_den, const real *v,
                  const real *jv, const real *rho)
{
    const size_t f = IX(i, k, j);
    if (!adv_den) return v[f] * jv[f];
    if (j == 0) return v[f] * jv[f] * (L(1.5) * rho[IX(i, k, 0)] - L(0.5) * rho[IX(i, k, 1)]);
    if (j == ny)
        return v[f] * jv[f] * (L(1.5) * rho[IX(i, k, ny - 1)] - L(0.5) * rho[IX(i, k, ny - 2)]);
    return v[f] * jv[f] * (rho[IX(i, k, j - 1)] + rho[IX(i, k, j)]) / L(2.0);
}

/* calc_divergence, wind.f90:201-246.  horz_only=1 is the form balance_uvw
 * uses; horz_only=0 adds the vertical mass-flux divergence with the
 * dz-weighted interface density (:229,236-244) and is used to check that a
 * balanced field has no divergence. */
void oracle_calc_divergence(int nx, int ny, int nz, real dx, int adv_den, int horz_only,
                            const real *u, const real *v, const real *w, const real *rho,
                            const real *ju, const real *jv, const real *jw, const real *dz,
                            real *div)
{
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                const real dU = u_met(nx, ny, nz, i + 1, k, j, adv_den, u, ju, rho) -
                                u_met(nx, ny, nz, i, k, j, adv_den, u, ju, rho);
                const real dV = v_met(nx, ny, nz, i, k, j + 1, adv_den, v, jv, rho) -
                                v_met(nx, ny, nz, i, k, j, adv_den, v, jv, rho);
                div[IX(i, k, j)] = (dU + dV) / dx;
            }
    if (horz_only) return;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            real below = 0;
            for (int k = 0; k < nz; k++) {
                const size_t c = IX(i, k, j);
                real wm;
                if (!adv_den)
                    wm = w[c] * jw[c];
                else if (k == nz - 1)
                    wm = w[c] * jw[c] * rho[c];
                else {
                    const size_t a = IX(i, k + 1, j);
                    const real rho_i = (rho[c] * dz[c] + rho[a] * dz[a]) / (dz[c] + dz[a]);
                    wm = w[c] * jw[c] * rho_i;
                }
                if (k == 0)
                    div[c] = div[c] + wm / dz[c];
                else
                    div[c] = div[c] + (wm - below) / dz[c];
                below = wm;
            }
        }
}

/* calc_w, wind.f90:166-199: w = 0 everywhere, then the columns
 * i_s..i_e x j_s..j_e (0-based, inclusive) integrated upward from the ground
 * so that the horizontal divergence div is balanced. */
void oracle_calc_w(int nx, int ny, int nz, int i_s, int i_e, int j_s, int j_e, int adv_den,
                   const real *div, const real *dz, const real *jw, const real *rho, real *w)
{
    for (size_t c = 0; c < (size_t)nx * ny * nz; c++) w[c] = 0;
    for (int j = j_s; j <= j_e; j++)
        for (int i = i_s; i <= i_e; i++)
            for (int k = 0; k < nz; k++) {
                const size_t c = IX(i, k, j);
                /* rho_i(k) = (rho(k)dz(k) + rho(k+1)dz(k+1)) / (dz(k)+dz(k+1)), :177 */
#define RHO_I(kk)                                                                            \
    ((rho[IX(i, kk, j)] * dz[IX(i, kk, j)] + rho[IX(i, (kk) + 1, j)] * dz[IX(i, (kk) + 1, j)]) / \
     (dz[IX(i, kk, j)] + dz[IX(i, (kk) + 1, j)]))
                if (adv_den) {
                    if (k == 0)
                        w[c] = L(0.0) - div[c] * dz[c] / (jw[c] * RHO_I(0));
                    else if (k == nz - 1) {
                        const size_t b = IX(i, k - 1, j);
                        w[c] = ((w[b] * RHO_I(k - 1) * jw[b]) - div[c] * dz[c]) / (jw[c] * rho[c]);
                    } else {
                        const size_t b = IX(i, k - 1, j);
                        w[c] = ((w[b] * RHO_I(k - 1) * jw[b]) - div[c] * dz[c]) / (jw[c] * RHO_I(k));
                    }
                } else {
                    if (k == 0)
                        w[c] = (L(0.0) - div[c] * dz[c]) / (jw[c]);
                    else {
                        const size_t b = IX(i, k - 1, j);
                        w[c] = (w[b] * jw[b] - div[c] * dz[c]) / (jw[c]);
                    }
                }
#undef RHO_I
            }
}

/* compute_dt, time_step.f90:51-168 with use_density=.false.: u, v, w the
 * winds (m/s), dz_levels the nz nominal level thicknesses, its..ite and
 * jts..jte the tile (0-based, inclusive).  Returns dt = CFL / maxwind3d in
 * the real type; the reference stops when dt < 0.1 (:158-166), which the
 * caller checks. */
real oracle_compute_dt(int nx, int ny, int nz, int its, int ite, int jts, int jte, real dx,
                       const real *u, const real *v, const real *w, const real *dz_levels,
                       real cfl, int strictness)
{
    const real three_d_cfl = L(0.577350269);
    const real sqrt3 = SQRT(L(3.0)) * L(1.001);
    real maxwind1d = 0, maxwind3d = 0;
    /* maxval(abs(.)) over the whole arrays */
    real mu = 0, mv = 0, mw = 0;
    for (size_t c = 0; c < (size_t)(nx + 1) * nz * ny; c++) mu = rmax(mu, fabs_r(u[c]));
    for (size_t c = 0; c < (size_t)nx * nz * (ny + 1); c++) mv = rmax(mv, fabs_r(v[c]));
    for (size_t c = 0; c < (size_t)nx * nz * ny; c++) mw = rmax(mw, fabs_r(w[c]));
    if (strictness == 1) {
        maxwind1d = rmax(mu, mv);
        maxwind1d = rmax(maxwind1d, mw);
        maxwind3d = maxwind1d * sqrt3;
    } else if (strictness == 5) {
        maxwind3d = mu + mv + mw;
    } else {
        for (int j = jts; j <= jte; j++)
            for (int k = 0; k < nz; k++) {
                const int zoff = k == 0 ? 0 : -1;
                for (int i = its; i <= ite; i++) {
                    const real a = rmax(fabs_r(u[IXU(i, k, j)]), fabs_r(u[IXU(i + 1, k, j)])) / dx;
                    const real b = rmax(fabs_r(v[IX(i, k, j)]), fabs_r(v[IX(i, k, j + 1)])) / dx;
                    const real c = rmax(fabs_r(w[IX(i, k, j)]), fabs_r(w[IX(i, k + zoff, j)])) /
                                   dz_levels[k];
                    const real cur = rmax(rmax(a, b), c);
                    maxwind3d = rmax(maxwind3d, cur);
                }
            }
        if (strictness == 2) {
            maxwind3d = maxwind3d * three_d_cfl;
            maxwind1d = rmax(mu, mv);
            maxwind1d = rmax(maxwind1d, mw);
            maxwind3d = rmax(maxwind1d, maxwind3d);
        } else if (strictness == 4) {
            maxwind3d = maxwind3d * sqrt3;
        }
    }
    return cfl / maxwind3d;
}

/* ------------------------------------------------------------------------- */
/* Lateral forcing of advected species (the step that fills the ring and     */
/* relaxes the border before advect in ICAR's time step)                     */
/* ------------------------------------------------------------------------- */

/* init_relax_filters, src/objects/domain_obj.f90:3285-3390: the 2-D filter
 * f[j*nx + i] (0-based) of a domain with halo hs and the given boundary
 * flags, statement by statement (later statements overwrite earlier ones;
 * a domain touching boundaries in both x and y takes the corner branches,
 * of which only the first matching pair applies). */
void oracle_relax_filter(int nx, int ny, int hs, int west, int east, int north, int south,
                         real *f)
{
    static const real rs[7] = {L(0.9), L(0.75), L(0.6), L(0.5), L(0.4), L(0.25), L(0.1)};
    static const real rs_r[7] = {L(0.1), L(0.25), L(0.4), L(0.5), L(0.6), L(0.75), L(0.9)};
#define F(i, j) f[(size_t)(j) * nx + (i)]
    int nr = 7;
    if (nx - 1 - hs < nr) nr = nx - 1 - hs; /* :3302 min(7, ime-ims-hs, jme-jms-hs) */
    if (ny - 1 - hs < nr) nr = ny - 1 - hs;
    for (size_t c = 0; c < (size_t)nx * ny; c++) f[c] = 0;
    const int corner = (west || east) && (north || south);
    if (!corner) {
        if (west) { /* :3309-3313 */
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < hs; i++) F(i, j) = 1;
            for (int j = 0; j < ny; j++)
                for (int r = 0; r < nr; r++) F(hs + r, j) = rs[r];
        } else if (east) { /* :3314-3318 */
            for (int j = 0; j < ny; j++)
                for (int i = nx - hs; i < nx; i++) F(i, j) = 1;
            for (int j = 0; j < ny; j++)
                for (int r = 0; r < nr; r++) F(nx - hs - nr + r, j) = rs_r[r];
        } else if (north) { /* :3319-3323, rows jme-hs-1:jme */
            for (int j = ny - hs - 2; j < ny; j++)
                for (int i = 0; i < nx; i++) F(i, j) = 1;
            for (int i = 0; i < nx; i++)
                for (int r = 0; r < nr; r++) F(i, ny - hs - nr + r) = rs_r[r];
        } else if (south) { /* :3324-3328 */
            for (int j = 0; j < hs; j++)
                for (int i = 0; i < nx; i++) F(i, j) = 1;
            for (int i = 0; i < nx; i++)
                for (int r = 0; r < nr; r++) F(i, hs + r) = rs[r];
        }
    } else if (north && west) { /* :3331-3345 */
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < hs; i++) F(i, j) = 1;
        for (int j = ny - hs - 2; j < ny; j++)
            for (int i = 0; i < nx; i++) F(i, j) = 1;
        for (int j = 0; j <= ny - 1 - hs; j++)
            for (int r = 0; r < nr; r++) F(hs + r, j) = rs[r];
        for (int i = hs; i < nx; i++)
            for (int r = 0; r < nr; r++) F(i, ny - hs - nr + r) = rs_r[r];
        for (int a = 1; a <= nr; a++)
            for (int b = 1; b <= nr; b++) F(hs + a - 1, ny - hs - b) = rs[(a < b ? a : b) - 1];
    } else if (north && east) { /* :3346-3360 */
        for (int j = 0; j < ny; j++)
            for (int i = nx - hs; i < nx; i++) F(i, j) = 1;
        for (int j = ny - hs - 2; j < ny; j++)
            for (int i = 0; i < nx; i++) F(i, j) = 1;
        for (int j = 0; j <= ny - 1 - hs; j++)
            for (int r = 0; r < nr; r++) F(nx - hs - nr + r, j) = rs_r[r];
        for (int i = 0; i <= nx - 1 - hs; i++)
            for (int r = 0; r < nr; r++) F(i, ny - hs - nr + r) = rs_r[r];
        for (int a = 1; a <= nr; a++)
            for (int b = 1; b <= nr; b++) F(nx - hs - a, ny - hs - b) = rs[(a < b ? a : b) - 1];
    } else if (south && west) { /* :3361-3375 */
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < hs; i++) F(i, j) = 1;
        for (int j = 0; j < hs; j++)
            for (int i = 0; i < nx; i++) F(i, j) = 1;
        for (int j = hs; j < ny; j++)
            for (int r = 0; r < nr; r++) F(hs + r, j) = rs[r];
        for (int i = hs; i < nx; i++)
            for (int r = 0; r < nr; r++) F(i, hs + r) = rs[r];
        for (int a = 1; a <= nr; a++)
            for (int b = 1; b <= nr; b++) F(hs + a - 1, hs + b - 1) = rs[(a < b ? a : b) - 1];
    } else if (south && east) { /* :3376-3390 */
        for (int j = 0; j < ny; j++)
            for (int i = nx - hs; i < nx; i++) F(i, j) = 1;
        for (int j = 0; j < hs; j++)
            for (int i = 0; i < nx; i++) F(i, j) = 1;
        for (int j = hs; j < ny; j++)
            for (int r = 0; r < nr; r++) F(nx - hs - nr + r, j) = rs_r[r];
        for (int i = 0; i <= nx - 1 - hs; i++)
            for (int r = 0; r < nr; r++) F(i, hs + r) = rs[r];
        for (int a = 1; a <= nr; a++)
            for (int b = 1; b <= nr; b++) F(nx - hs - a, hs + b - 1) = rs[(a < b ? a : b) - 1];
    }
#undef F
}

/* apply_forcing for a 3-D variable forced on the boundaries,
 * domain_obj.f90:3493-3503.  dt%seconds() is double precision, so the
 * tendency and the relaxation term are formed in double and the sums
 * rounded to the real kind on assignment; forcing - value is a difference
 * of two reals. */
void oracle_apply_forcing(int nx, int ny, int nz, double dt, const real *filt, real *q,
                          real *fhi, const real *dqdt)
{
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                const size_t c = IX(i, k, j);
                fhi[c] = (real)((double)fhi[c] + ((double)dqdt[c] * dt));
                const real rf = filt[(size_t)j * nx + i];
                if (rf == L(1.0))
                    q[c] = fhi[c];
                else if (rf < L(1.0))
                    q[c] = (real)((double)q[c] + ((double)rf * dt / 3600.0) * (double)(fhi[c] - q[c]));
            }
}
