/*
 * CPU restatement of ICAR's standard advection scheme (adv=1): flux-form
 * upwind / 3rd / 5th order fluxes, optional RK3 time stepping and the WRF
 * monotonic flux limiter.
 *
 * TEST INFRASTRUCTURE ONLY.  The test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product library never links
 * it and has no CPU fallback.
 *
 * It restates, in plain C with the same operation order, the Fortran in
 *   src/physics/advect.f90          adv_std_init (:29-71, the flux bounds),
 *                                   flux3 (:134-257), adv_std_advect3d
 *                                   (:259-320), adv_std_compute_wind
 *                                   (:444-492)
 *   src/physics/adv_fluxcorr.f90    WRF_flux_corr (:37-140), upwind_flux3
 *                                   (:142-190)
 *   src/physics/advection_driver.f90:96-119  the RK3 sequence and the Euler
 *                                   step per species
 * Every Fortran expression is evaluated left to right as written, without
 * contraction (-ffp-contract=off).  Fortran MAX/MIN of two reals are
 * a > b ? a : b and a < b ? a : b (the first argument on ties, so the sign
 * of a zero result follows that rule); the HIP library uses the same
 * selects.  -DORACLE_DOUBLE=1 gives the 8-byte-real build.
 *
 * Indices are 0-based memory offsets (i - ims, k - kms, j - jms).  Layouts:
 *   cells     IX(i,k,j)  = i + nx*(k + nz*j)
 *   x faces   FXI(i,k,j) = i + (nx+1)*(k + nz*j), face i on the west side
 *             of cell i (the layout of domain%u)
 *   y faces   FYI(i,k,j) = i + nx*(k + nz*j) for j = 0..ny, face j on the
 *             south side of row j (the layout of domain%v)
 *   z faces   FZI(i,k,j) = i + nx*(k + (nz+1)*j), face k under level k,
 *             k = nz the model top
 * U_m is stored on x faces (faces 1..nx-1 set, as U_m(ims+1:ime)), V_m on
 * y faces (rows 1..ny-1), W_m on cells (w of the top face of the cell).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#if ORACLE_DOUBLE
typedef double real;
#define L(x) (x)
#define FABS fabs
#else
typedef float real;
#define L(x) (x##f)
#define FABS fabsf
#endif

#define IX(i, k, j) ((size_t)(i) + (size_t)nx * ((size_t)(k) + (size_t)nz * (size_t)(j)))
#define FXI(i, k, j) ((size_t)(i) + (size_t)(nx + 1) * ((size_t)(k) + (size_t)nz * (size_t)(j)))
#define FYI(i, k, j) IX(i, k, j)
#define FZI(i, k, j) ((size_t)(i) + (size_t)nx * ((size_t)(k) + (size_t)(nz + 1) * (size_t)(j)))

static inline real fmx(real a, real b) { return a > b ? a : b; }
static inline real fmn(real a, real b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------- */
/* adv_std_compute_wind, advect.f90:444-492                                   */
/* ------------------------------------------------------------------------- */
void oracle_std_compute_wind(int nx, int ny, int nz, real dx, real dt, int advect_density,
                             const real *u, const real *v, const real *w, const real *density,
                             const real *adz, const real *jaco_u, const real *jaco_v,
                             const real *jaco_w, real *Um, real *Vm, real *Wm, real *rho)
{
    const size_t N = (size_t)nx * nz * ny;
    /* :472-473  rho = 1, or the density */
    for (size_t c = 0; c < N; c++) rho[c] = advect_density ? density[c] : L(1.0);
    memset(Um, 0, sizeof(real) * (size_t)(nx + 1) * nz * ny);
    memset(Vm, 0, sizeof(real) * (size_t)nx * nz * (ny + 1));
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 1; i < nx; i++) {  /* :478-479, U_m(ims+1:ime) */
                const size_t f = FXI(i, k, j);
                Um[f] = u[f] * dt * (rho[IX(i, k, j)] + rho[IX(i - 1, k, j)]) * L(0.5) * jaco_u[f] / dx;
            }
    for (int j = 1; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {  /* :480-481, V_m(:,:,jms+1:jme) */
                const size_t f = FYI(i, k, j);
                Vm[f] = v[f] * dt * (rho[IX(i, k, j)] + rho[IX(i, k, j - 1)]) * L(0.5) * jaco_v[f] / dx;
            }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                const size_t c = IX(i, k, j);
                if (k < nz - 1) {  /* :475-476 rho_i, :483 */
                    const size_t a = IX(i, k + 1, j);
                    const real ri = (rho[c] * adz[c] + rho[a] * adz[a]) / (adz[c] + adz[a]);
                    Wm[c] = w[c] * dt * jaco_w[c] * ri;
                } else {  /* :484 */
                    Wm[c] = w[c] * dt * jaco_w[c] * rho[c];
                }
            }
}

/* ------------------------------------------------------------------------- */
/* flux3, advect.f90:134-257: one face.  m3..p2 are q at offsets -3..+2 from */
/* the cell on the positive side of the face (p0 = that cell, m1 = the one   */
/* on the negative side).                                                    */
/* ------------------------------------------------------------------------- */
static inline real face_upwind(real u, real m1, real p0)
{
    return ((u + FABS(u)) * m1 + (u - FABS(u)) * p0) / L(2.0);  /* :150-151 */
}

static inline real face_o3(real u, real m2, real m1, real p0, real p1)
{
    real f = L(7.0) * (p0 + m1) - (p1 + m2);                         /* :154-155 */
    f = (u * f) / L(12.0);                                            /* :156 */
    f = f - (FABS(u) / L(12.0)) * (L(3.0) * (p0 - m1) - (p1 - m2));   /* :158-159 */
    return f;
}

static inline real face_o5(real u, real m3, real m2, real m1, real p0, real p1, real p2)
{
    real f = L(37.0) * (p0 + m1) - L(8.0) * (p1 + m2) + (p2 + m3);   /* :171-173 */
    f = (u * f) / L(60.0);                                            /* :174 */
    f = f - (FABS(u) / L(60.0)) *
                (L(10.0) * (p0 - m1) - L(5.0) * (p1 - m2) + (p2 - m3)); /* :176-178 */
    return f;
}

static inline real face_h(int order, real u, const real *q, ptrdiff_t s)
{
    /* q points at the cell on the positive side; s is the stride */
    if (order == 1) return face_upwind(u, q[-s], q[0]);
    if (order == 3) return face_o3(u, q[-2 * s], q[-s], q[0], q[s]);
    return face_o5(u, q[-3 * s], q[-2 * s], q[-s], q[0], q[s], q[2 * s]);
}

/* vertical face kf (1..nz-1) of a column whose levels are q[k*s] */
static inline real face_v(int order, int nz, int kf, real w, const real *q, ptrdiff_t s)
{
    const real *p = q + (ptrdiff_t)kf * s;
    /* the upwind faces next to the ground and the top are written last
     * (:183-185, :198-201, :237-240) */
    if (order == 1 || kf == 1 || kf == nz - 1) return face_upwind(w, p[-s], p[0]);
    if (order == 5 && kf >= 3 && kf <= nz - 3)  /* :208-217 */
        return face_o5(w, p[-3 * s], p[-2 * s], p[-s], p[0], p[s], p[2 * s]);
    return face_o3(w, p[-2 * s], p[-s], p[0], p[s]);  /* :190-197, :219-233 */
}

/* flux3 over cells [i_s,i_e] x [j_s,j_e]; u, v, w are the t_factor-scaled
 * winds in the U_m / V_m / W_m layouts. */
static void flux3(int nx, int ny, int nz, int i_s, int i_e, int j_s, int j_e, int horder,
                  int vorder, const real *q, const real *u, const real *v, const real *w,
                  real *fx, real *fy, real *fz)
{
    (void)ny;
    for (int j = j_s; j <= j_e; j++)
        for (int k = 0; k < nz; k++)
            for (int i = i_s; i <= i_e + 1; i++)
                fx[FXI(i, k, j)] = face_h(horder, u[FXI(i, k, j)], q + IX(i, k, j), 1);
    for (int j = j_s; j <= j_e + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = i_s; i <= i_e; i++)
                fy[FYI(i, k, j)] = face_h(horder, v[FYI(i, k, j)], q + IX(i, k, j),
                                          (ptrdiff_t)nx * nz);
    for (int j = j_s; j <= j_e; j++)
        for (int i = i_s; i <= i_e; i++) {
            const real *col = q + IX(i, 0, j);
            for (int kf = 1; kf < nz; kf++)
                fz[FZI(i, kf, j)] = face_v(vorder, nz, kf, w[IX(i, kf - 1, j)], col, nx);
            fz[FZI(i, 0, j)] = L(0.0);                                               /* :252 */
            fz[FZI(i, nz, j)] = q[IX(i, nz - 1, j)] * w[IX(i, nz - 1, j)];           /* :253 */
        }
}

/* ------------------------------------------------------------------------- */
/* upwind_flux3, adv_fluxcorr.f90:142-190: two half-step upwind passes over  */
/* the extended cells [its-1,ite+1] x [jts-1,jte+1].                         */
/* ------------------------------------------------------------------------- */
static inline real half_upwind(real u, real m1, real p0)
{
    return L(0.5) * ((u + FABS(u)) * m1 + (u - FABS(u)) * p0) / L(2.0);
}

static void upwind_flux3(int nx, int ny, int nz, int its, int ite, int jts, int jte,
                         const real *q, const real *u, const real *v, const real *w,
                         const real *jaco, const real *dz, const real *rho, real *fx, real *fy,
                         real *fz, real *dq)
{
    const size_t N = (size_t)nx * nz * ny;
    const size_t sj = (size_t)nx * nz;
    for (int pass = 0; pass < 2; pass++) {
        const real *s = pass == 0 ? q : dq;
        for (int j = jts - 1; j <= jte + 1; j++)
            for (int k = 0; k < nz; k++)
                for (int i = its - 1; i <= ite + 2; i++) {  /* :160, :176 */
                    const size_t f = FXI(i, k, j);
                    const real t = half_upwind(u[f], s[IX(i - 1, k, j)], s[IX(i, k, j)]);
                    fx[f] = pass == 0 ? t : fx[f] + t;
                }
        for (int j = jts - 1; j <= jte + 2; j++)
            for (int k = 0; k < nz; k++)
                for (int i = its - 1; i <= ite + 1; i++) {  /* :162, :178 */
                    const size_t f = FYI(i, k, j);
                    const real t = half_upwind(v[f], s[IX(i, k, j) - sj], s[IX(i, k, j)]);
                    fy[f] = pass == 0 ? t : fy[f] + t;
                }
        for (int j = jts - 1; j <= jte + 1; j++)
            for (int i = its - 1; i <= ite + 1; i++) {
                for (int kf = 1; kf < nz; kf++) {  /* :164-165, :180-181 */
                    const size_t f = FZI(i, kf, j);
                    const real t = half_upwind(w[IX(i, kf - 1, j)], s[IX(i, kf - 1, j)], s[IX(i, kf, j)]);
                    fz[f] = pass == 0 ? t : fz[f] + t;
                }
                const size_t top = FZI(i, nz, j);
                const real t = L(0.5) * s[IX(i, nz - 1, j)] * w[IX(i, nz - 1, j)]; /* :169, :184 */
                if (pass == 0) {
                    fz[FZI(i, 0, j)] = L(0.0);                                  /* :168 */
                    fz[top] = t;
                } else {
                    fz[top] = fz[top] + t;
                }
            }
        if (pass == 0) {  /* :172-175 dumb_q */
            memcpy(dq, q, N * sizeof(real));
            for (int j = jts - 1; j <= jte + 1; j++)
                for (int k = 0; k < nz; k++)
                    for (int i = its - 1; i <= ite + 1; i++) {
                        const size_t c = IX(i, k, j);
                        real t = q[c] - ((fx[FXI(i + 1, k, j)] - fx[FXI(i, k, j)]) +
                                         (fy[FYI(i, k, j + 1)] - fy[FYI(i, k, j)])) /
                                            (jaco[c] * rho[c]);
                        t = t - (fz[FZI(i, k + 1, j)] - fz[FZI(i, k, j)]) / (dz[c] * jaco[c] * rho[c]);
                        dq[c] = t;
                    }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* WRF_flux_corr, adv_fluxcorr.f90:37-140 (Wang et al. 2009): limits the     */
/* high-order fluxes fx, fy, fz in place over the extended cells.            */
/* ------------------------------------------------------------------------- */
static void wrf_flux_corr(int nx, int ny, int nz, int its, int ite, int jts, int jte,
                          const real *q, const real *u, const real *v, const real *w,
                          real *fx, real *fy, real *fz, const real *jaco, const real *dz,
                          const real *rho)
{
    const size_t N = (size_t)nx * nz * ny;
    const size_t NX = (size_t)(nx + 1) * nz * ny, NY = (size_t)nx * nz * (ny + 1),
                 NZ = (size_t)nx * (nz + 1) * ny;
    const size_t sj = (size_t)nx * nz;
    real *ufx = calloc(NX, sizeof(real)), *ufy = calloc(NY, sizeof(real)),
         *ufz = calloc(NZ, sizeof(real)), *dq = malloc(N * sizeof(real));
    real *sin_ = malloc(N * sizeof(real)), *sout = malloc(N * sizeof(real));

    upwind_flux3(nx, ny, nz, its, ite, jts, jte, q, u, v, w, jaco, dz, rho, ufx, ufy, ufz, dq);

    /* :80-85 merged fluxes */
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 2; i++) fx[FXI(i, k, j)] = fx[FXI(i, k, j)] - ufx[FXI(i, k, j)];
    for (int j = jts - 1; j <= jte + 2; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) fy[FYI(i, k, j)] = fy[FYI(i, k, j)] - ufy[FYI(i, k, j)];
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k <= nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) fz[FZI(i, k, j)] = fz[FZI(i, k, j)] - ufz[FZI(i, k, j)];

    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) {
                const size_t c = IX(i, k, j);
                const real jr = jaco[c] * rho[c];
                const real djr = dz[c] * jaco[c] * rho[c];
                /* :75-78 temp, the upwind-only value */
                real temp = q[c] - ((ufx[FXI(i + 1, k, j)] - ufx[FXI(i, k, j)]) +
                                    (ufy[FYI(i, k, j + 1)] - ufy[FYI(i, k, j)])) / jr;
                temp = temp - (ufz[FZI(i, k + 1, j)] - ufz[FZI(i, k, j)]) / djr;
                const real xe = fx[FXI(i + 1, k, j)], xw = fx[FXI(i, k, j)];
                const real yn = fy[FYI(i, k, j + 1)], ys = fy[FYI(i, k, j)];
                const real zt = fz[FZI(i, k + 1, j)], zb = fz[FZI(i, k, j)];
                /* :88-96 */
                real fin = -((fmn(L(0.0), xe) - fmx(L(0.0), xw)) + (fmn(L(0.0), yn) - fmx(L(0.0), ys))) / jr;
                fin = fin - (fmn(L(0.0), zt) - fmx(L(0.0), zb)) / djr;
                real fout = ((fmx(L(0.0), xe) - fmn(L(0.0), xw)) + (fmx(L(0.0), yn) - fmn(L(0.0), ys))) / jr;
                fout = fout + (fmx(L(0.0), zt) - fmn(L(0.0), zb)) / djr;
                /* :99-121 bounds from the upwind neighbours */
                real qmax = q[c], qmin = q[c];
                if (u[FXI(i, k, j)] > L(0.0)) {
                    qmax = fmx(q[c - 1], qmax);
                    qmin = fmn(q[c - 1], qmin);
                } else if (u[FXI(i + 1, k, j)] < L(0.0)) {
                    qmax = fmx(q[c + 1], qmax);
                    qmin = fmn(q[c + 1], qmin);
                }
                if (v[FYI(i, k, j)] > L(0.0)) {
                    qmax = fmx(q[c - sj], qmax);
                    qmin = fmn(q[c - sj], qmin);
                } else if (v[FYI(i, k, j + 1)] < L(0.0)) {
                    qmax = fmx(q[c + sj], qmax);
                    qmin = fmn(q[c + sj], qmin);
                }
                /* :115-121: w of the face below lifts from the level below;
                 * w of the face above, when negative, from the level above.
                 * Both are maxima, so their order does not matter. */
                if (k > 0 && w[c - nx] > L(0.0)) {
                    qmax = fmx(q[c - nx], qmax);
                    qmin = fmn(q[c - nx], qmin);
                }
                if (k < nz - 1 && w[c] < L(0.0)) {
                    qmax = fmx(q[c + nx], qmax);
                    qmin = fmn(q[c + nx], qmin);
                }
                /* :124-125 */
                real si = L(1.0), so = L(1.0);
                if (fin > (qmax - temp)) si = fmx(L(0.0), (qmax - temp) / (fin + L(1.E-15)));
                if (fout > (temp - qmin)) so = fmx(L(0.0), (temp - qmin) / (fout + L(1.E-15)));
                sin_[c] = si;
                sout[c] = so;
            }

    /* :129-145 normalised fluxes on the tile faces */
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its; i <= ite + 1; i++) {
                const size_t f = FXI(i, k, j), c = IX(i, k, j);
                if (fx[f] > L(0.0))
                    fx[f] = fmn(sin_[c], sout[c - 1]) * fx[f];
                else if (fx[f] < L(0.0))
                    fx[f] = fmn(sout[c], sin_[c - 1]) * fx[f];
            }
    for (int j = jts; j <= jte + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) {
                const size_t f = FYI(i, k, j), c = IX(i, k, j);
                if (fy[f] > L(0.0))
                    fy[f] = fmn(sin_[c], sout[c - sj]) * fy[f];
                else if (fy[f] < L(0.0))
                    fy[f] = fmn(sout[c], sin_[c - sj]) * fy[f];
            }
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 1; k < nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) {
                const size_t f = FZI(i, k, j), c = IX(i, k, j);
                if (fz[f] > L(0.0))
                    fz[f] = fmn(sin_[c], sout[c - nx]) * fz[f];
                else if (fz[f] < L(0.0))
                    fz[f] = fmn(sout[c], sin_[c - nx]) * fz[f];
            }

    /* :148-150 */
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 2; i++) fx[FXI(i, k, j)] = fx[FXI(i, k, j)] + ufx[FXI(i, k, j)];
    for (int j = jts - 1; j <= jte + 2; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) fy[FYI(i, k, j)] = fy[FYI(i, k, j)] + ufy[FYI(i, k, j)];
    for (int j = jts - 1; j <= jte + 1; j++)
        for (int k = 0; k <= nz; k++)
            for (int i = its - 1; i <= ite + 1; i++) fz[FZI(i, k, j)] = fz[FZI(i, k, j)] + ufz[FZI(i, k, j)];

    free(ufx);
    free(ufy);
    free(ufz);
    free(dq);
    free(sin_);
    free(sout);
}

/* ------------------------------------------------------------------------- */
/* adv_std_advect3d, advect.f90:259-320.  qf is qfluxes (in: the state the   */
/* fluxes are computed from; out: the new state), qold the base state; they  */
/* may alias (the Euler call of advection_driver.f90:118).  flux_corr 1 is   */
/* kFLUXCOR_MONO; fc_bounds is the module-level option that widens the flux  */
/* bounds by one cell (adv_std_init, advect.f90:57-62).                      */
/* ------------------------------------------------------------------------- */
void oracle_std_advect3d(int nx, int ny, int nz, int its, int ite, int jts, int jte, int horder,
                         int vorder, int fc_bounds, real *qf, const real *qold, const real *dz,
                         const real *jaco, const real *Um, const real *Vm, const real *Wm,
                         const real *rho, real t_factor, int flux_corr)
{
    const size_t N = (size_t)nx * nz * ny;
    const size_t NX = (size_t)(nx + 1) * nz * ny, NY = (size_t)nx * nz * (ny + 1),
                 NZ = (size_t)nx * (nz + 1) * ny;
    const int e = fc_bounds ? 1 : 0;
    const int i_s = its - e, i_e = ite + e, j_s = jts - e, j_e = jte + e;
    real *fx = calloc(NX, sizeof(real)), *fy = calloc(NY, sizeof(real)),
         *fz = calloc(NZ, sizeof(real));
    real *ut = malloc(NX * sizeof(real)), *vt = malloc(NY * sizeof(real)), *wt = malloc(N * sizeof(real));
    /* :285-287 the winds times t_factor (array temporaries) */
    for (size_t f = 0; f < NX; f++) ut[f] = Um[f] * t_factor;
    for (size_t f = 0; f < NY; f++) vt[f] = Vm[f] * t_factor;
    for (size_t c = 0; c < N; c++) wt[c] = Wm[c] * t_factor;
    flux3(nx, ny, nz, i_s, i_e, j_s, j_e, horder, vorder, qf, ut, vt, wt, fx, fy, fz);
    if (flux_corr == 1) /* :289-297, unscaled winds */
        wrf_flux_corr(nx, ny, nz, its, ite, jts, jte, qold, Um, Vm, Wm, fx, fy, fz, jaco, dz, rho);
    if (qf != qold) memcpy(qf, qold, N * sizeof(real)); /* :298 */
    for (int j = jts; j <= jte; j++)
        for (int k = 0; k < nz; k++)
            for (int i = its; i <= ite; i++) {
                const size_t c = IX(i, k, j);
                /* :301-304 */
                real t = qf[c] - ((fx[FXI(i + 1, k, j)] - fx[FXI(i, k, j)]) +
                                  (fy[FYI(i, k, j + 1)] - fy[FYI(i, k, j)])) /
                                     (jaco[c] * rho[c]);
                /* :307-309 */
                t = t - (fz[FZI(i, k + 1, j)] - fz[FZI(i, k, j)]) / (dz[c] * jaco[c] * rho[c]);
                qf[c] = t;
            }
    free(fx);
    free(fy);
    free(fz);
    free(ut);
    free(vt);
    free(wt);
}

/* advection_driver.f90:96-119 for one species: RK3 (three calls with
 * t_factor 0.333, 0.5 and 1, the last with the flux correction, each from
 * the original state) or one Euler call. */
void oracle_std_advect_species(int nx, int ny, int nz, int its, int ite, int jts, int jte,
                               int horder, int vorder, int flux_corr, int rk3, real *q,
                               const real *dz, const real *jaco, const real *Um, const real *Vm,
                               const real *Wm, const real *rho)
{
    const size_t N = (size_t)nx * nz * ny;
    const int fcb = flux_corr == 1;
    if (rk3) {
        real *temp = malloc(N * sizeof(real));
        memcpy(temp, q, N * sizeof(real));  /* :103 */
        oracle_std_advect3d(nx, ny, nz, its, ite, jts, jte, horder, vorder, fcb, temp, q, dz, jaco,
                            Um, Vm, Wm, rho, L(0.333), 0);
        oracle_std_advect3d(nx, ny, nz, its, ite, jts, jte, horder, vorder, fcb, temp, q, dz, jaco,
                            Um, Vm, Wm, rho, L(0.5), 0);
        oracle_std_advect3d(nx, ny, nz, its, ite, jts, jte, horder, vorder, fcb, temp, q, dz, jaco,
                            Um, Vm, Wm, rho, L(1.0), flux_corr);
        memcpy(q, temp, N * sizeof(real));  /* :110 */
        free(temp);
    } else {
        oracle_std_advect3d(nx, ny, nz, its, ite, jts, jte, horder, vorder, fcb, q, q, dz, jaco,
                            Um, Vm, Wm, rho, L(1.0), flux_corr);
    }
}

/* Exported for the pinning tests: flux3 (advect.f90:134-257) over cells
 * [i_s,i_e] x [j_s,j_e] with the given (already t_factor-scaled) winds, and
 * WRF_flux_corr (adv_fluxcorr.f90:37-140) applied to given fluxes. */
void oracle_std_flux3(int nx, int ny, int nz, int i_s, int i_e, int j_s, int j_e, int horder,
                      int vorder, const real *q, const real *u, const real *v, const real *w,
                      real *fx, real *fy, real *fz)
{
    flux3(nx, ny, nz, i_s, i_e, j_s, j_e, horder, vorder, q, u, v, w, fx, fy, fz);
}

void oracle_std_flux_corr(int nx, int ny, int nz, int its, int ite, int jts, int jte,
                          const real *q, const real *u, const real *v, const real *w, real *fx,
                          real *fy, real *fz, const real *jaco, const real *dz, const real *rho)
{
    wrf_flux_corr(nx, ny, nz, its, ite, jts, jte, q, u, v, w, fx, fy, fz, jaco, dz, rho);
}
