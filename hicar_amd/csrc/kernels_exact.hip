// "Exact" MPDATA kernels: one thread per cell, each expression evaluated in
// the order the Fortran writes it, no FMA contraction (this translation unit
// is compiled with -ffp-contract=off) and IEEE-correct FP32 division, so an
// FP32 step is bit-identical to a strict IEEE FP32 build of the reference.
// These kernels are the parity baseline; the fused kernels of
// kernels_fast.hip are checked against them and against the oracle.
//
// Reference: src/physics/adv_mpdata.f90 and adv_mpdata_FCT_core.f90.
#include "common.hpp"

namespace icaradv {

namespace {

constexpr int BX = 64;  // one wavefront along i (coalesced rows)
constexpr int BY = 4;   // rows j per block

__device__ __forceinline__ float flux1(float l, float r, float U)
{
    // adv_mpdata.f90:39  f = ((U+|U|)*l + (U-|U|)*r)/2
    return ((U + fabsf(U)) * l + (U - fabsf(U)) * r) / 2.0f;
}

__device__ __forceinline__ bool cell_coords(const Dims &d, int &i, int &k, int &j)
{
    i = blockIdx.x * BX + threadIdx.x;
    j = blockIdx.z * BY + threadIdx.y;
    k = blockIdx.y;
    return i < d.nx && j < d.ny;
}

dim3 cell_grid(const Dims &d) { return dim3((d.nx + BX - 1) / BX, d.nz, (d.ny + BY - 1) / BY); }

// ---------------------------------------------------------------------------
// mpdata_compute_wind (adv_mpdata.f90:477-504) plus the species-invariant
// products mpdata_advect3d forms per species (:374 W_m/dz, jaco*rho) and the
// denominators of upwind_advection (:87, :91-97).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BX *BY)
    k_courant(Dims d, float dx, float dt, int advect_density, DomainFields f, StepFields o)
{
    int i, k, j;
    if (!cell_coords(d, i, k, j)) return;
    const size_t c = ix(d, i, k, j);
    const size_t sj = (size_t)d.nx * d.nz;
    const float rc = advect_density ? f.density[c] : 1.0f;  // :489-490
    if (i >= 1) {
        const size_t iu = (size_t)i + (size_t)(d.nx + 1) * ((size_t)k + (size_t)d.nz * j);
        const float rl = advect_density ? f.density[c - 1] : 1.0f;
        o.U[c] = f.u[iu] * dt * (rc + rl) * 0.5f * f.jaco_u[iu] / dx;  // :492-493
    } else {
        o.U[c] = 0.0f;
    }
    if (j >= 1) {
        const float rs = advect_density ? f.density[c - sj] : 1.0f;
        o.V[c] = f.v[c] * dt * (rc + rs) * 0.5f * f.jaco_v[c] / dx;  // :494-495
    } else {
        o.V[c] = 0.0f;
    }
    float Wc;
    if (k < d.nz - 1) {
        const float ru = advect_density ? f.density[c + d.nx] : 1.0f;
        Wc = f.w[c] * dt * f.jaco_w[c] * (ru + rc) * 0.5f;  // :496-497
    } else {
        Wc = f.w[c] * dt * f.jaco_w[c] * rc;  // :498
    }
    o.W[c] = Wc;
    const float jac = f.jaco[c], den = f.density[c], dzc = f.dz[c];
    o.G[c] = jac * den;          // jaco*rho with rho = domain density (:374, :87)
    o.Gz[c] = dzc * jac * den;   // dz*jaco*rho (:91,94,97)
    o.Wd[c] = Wc / dzc;          // W_m/dz (:374)
}

// ---------------------------------------------------------------------------
// upwind_advection (adv_mpdata.f90:43-103): donor cell, ring copied.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BX *BY)
    k_upwind(Dims d, const float *__restrict__ qin, const float *__restrict__ U,
             const float *__restrict__ V, const float *__restrict__ W, const float *__restrict__ G,
             const float *__restrict__ Gz, float *__restrict__ q)
{
    int i, k, j;
    if (!cell_coords(d, i, k, j)) return;
    const size_t c = ix(d, i, k, j);
    const float qc = qin[c];
    if (i == 0 || i == d.nx - 1 || j == 0 || j == d.ny - 1) {  // :61-63
        q[c] = qc;
        return;
    }
    const size_t sj = (size_t)d.nx * d.nz;
    const float f1r = flux1(qc, qin[c + 1], U[c + 1]);       // :72-73
    const float f1l = flux1(qin[c - 1], qc, U[c]);
    const float f3 = flux1(qc, qin[c + sj], V[c + sj]);      // :75-76
    const float f4 = flux1(qin[c - sj], qc, V[c]);           // :78-79
    float qn = qc - ((f1r - f1l) + (f3 - f4)) / G[c];         // :86-87
    const float den = Gz[c];
    if (k == 0) {
        const float f5 = flux1(qc, qin[c + d.nx], W[c]);
        qn = qn - f5 / den;                                    // :93-94
    } else if (k == d.nz - 1) {
        const float f5m = flux1(qin[c - d.nx], qc, W[c - d.nx]);
        qn = qn - (qc * W[c] - f5m) / den;                     // :96-97
    } else {
        const float f5 = flux1(qc, qin[c + d.nx], W[c]);
        const float f5m = flux1(qin[c - d.nx], qc, W[c - d.nx]);
        qn = qn - (f5 - f5m) / den;                            // :90-91
    }
    q[c] = qn;
}

// ---------------------------------------------------------------------------
// mpdata_fluxes (adv_mpdata.f90:105-252) with the scaling of :378-380.
// Each thread writes the x face on the left of its cell, the y face on its
// south side and the z face on its top.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BX *BY)
    k_pseudo(Dims d, const float *__restrict__ q, const float *__restrict__ U,
             const float *__restrict__ V, const float *__restrict__ Wd, const float *__restrict__ G,
             const float *__restrict__ dz, float *__restrict__ u2, float *__restrict__ v2,
             float *__restrict__ w2)
{
    int i, k, j;
    if (!cell_coords(d, i, k, j)) return;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const size_t c = ix(d, i, k, j);
    const size_t sk = (size_t)nx, sj = (size_t)nx * nz;
    const float eps = 1e-10f;
    auto Q = [&](int di, int dk, int dj) {
        return q[c + (ptrdiff_t)di + (ptrdiff_t)dk * (ptrdiff_t)sk + (ptrdiff_t)dj * (ptrdiff_t)sj];
    };

    // ---- U component (:131-165)
    if (i >= 1) {
        const size_t cl = c - 1;
        const float rx = q[c], lx = q[cl];
        const float denom = rx + lx + eps;
        const float Uc = U[c];
        const float au = fabsf(Uc);
        float val = au * (1.0f - au / (0.5f * (G[c] + G[cl])));
        val = val * (rx - lx) / denom;
        if (j > 0 && j < ny - 1) {  // UxV (:145-152)
            const float eq = (Q(0, 0, 1) - Q(0, 0, -1) + Q(-1, 0, 1) - Q(-1, 0, -1)) /
                             (Q(0, 0, 1) + Q(0, 0, -1) + Q(-1, 0, 1) + Q(-1, 0, -1) + eps);
            const float ev = (1.0f / 4.0f) * (V[c] + V[c + sj] + V[c - 1] + V[c - 1 + sj]);
            val = val - 0.5f * Uc * ev * eq / (G[c] + G[cl]);
        }
        if (k > 0 && k < nz - 1) {  // UxW (:157-164)
            const float eq = (Q(0, 1, 0) - Q(0, -1, 0) + Q(-1, 1, 0) - Q(-1, -1, 0)) /
                             (Q(0, 1, 0) + Q(0, -1, 0) + Q(-1, 1, 0) + Q(-1, -1, 0) + eps);
            const float ev = (1.0f / 4.0f) * (Wd[c] + Wd[c - sk] + Wd[c - 1] + Wd[c - 1 - sk]);
            val = val - 0.5f * Uc * ev * eq / (G[c] + G[cl]);
        }
        u2[c] = val * 0.5f;  // :378
    } else {
        u2[c] = 0.0f;
    }

    // ---- V component (:169-205)
    if (j >= 1) {
        const size_t cs = c - sj;
        const float r = q[c], l = q[cs];
        const float denom = r + l + eps;
        const float Vc = V[c];
        const float av = fabsf(Vc);
        float val = av * (1.0f - av / (0.5f * (G[c] + G[cs])));
        val = val * (r - l) / denom;
        if (i > 0 && i < nx - 1) {  // VxU (:186-192); zero edge terms at i = 0, nx-1
            const float eq = (Q(1, 0, -1) - Q(-1, 0, 0) + Q(1, 0, 0) - Q(-1, 0, -1)) /
                             (Q(1, 0, 0) + Q(1, 0, -1) + Q(-1, 0, 0) + Q(-1, 0, -1) + eps);
            const float ev = (1.0f / 4.0f) * (U[c + 1] + U[c + 1 - sj] + U[c] + U[c - sj]);
            val = val - 0.5f * Vc * ev * eq / (G[c] + G[cs]);
        }
        if (k > 0 && k < nz - 1) {  // VxW (:197-204)
            const float eq = (Q(0, 1, -1) - Q(0, -1, 0) + Q(0, 1, 0) - Q(0, -1, -1)) /
                             (Q(0, 1, -1) + Q(0, -1, 0) + Q(0, 1, 0) + Q(0, -1, -1) + eps);
            const float ev = (1.0f / 4.0f) * (Wd[c] + Wd[c - sk] + Wd[c - sj] + Wd[c - sk - sj]);
            val = val - 0.5f * Vc * ev * eq / (G[c] + G[cs]);
        }
        v2[c] = val * 0.5f;  // :379
    } else {
        v2[c] = 0.0f;
    }

    // ---- W component (:211-246)
    if (k < nz - 1) {
        const size_t cu = c + sk;
        const float r = q[cu], l = q[c];
        const float denom = r + l + eps;
        const float Wc = Wd[c];
        const float aw = fabsf(Wc);
        float val = aw * (1.0f - aw / (0.5f * (G[cu] + G[c])));
        val = val * (r - l) / denom;
        if (i > 0 && i < nx - 1) {  // WxU (:227-233)
            const float eq = (Q(1, 1, 0) - Q(-1, 0, 0) + Q(1, 0, 0) - Q(-1, 1, 0)) /
                             (Q(1, 0, 0) + Q(1, 1, 0) + Q(-1, 0, 0) + Q(-1, 1, 0) + eps);
            const float ev = (1.0f / 4.0f) * (U[c + 1] + U[c + 1 + sk] + U[c] + U[c + sk]);
            val = val - 0.5f * Wc * ev * eq / (G[cu] + G[c]);
        }
        if (j > 0 && j < ny - 1) {  // WxV (:238-245)
            const float eq = (Q(0, 1, 1) - Q(0, 0, -1) + Q(0, 0, 1) - Q(0, 1, -1)) /
                             (Q(0, 0, 1) + Q(0, 1, -1) + Q(0, 1, 1) + Q(0, 0, -1) + eps);
            const float ev = (1.0f / 4.0f) * (V[c] + V[c + sk] + V[c + sj] + V[c + sk + sj]);
            val = val - 0.5f * Wc * ev * eq / (G[cu] + G[c]);
        }
        w2[c] = val * 0.5f * dz[c];  // :380
    } else {
        w2[c] = 0.0f;  // :211-212
    }
}

// ---------------------------------------------------------------------------
// The 1-D FCT of adv_mpdata_FCT_core.f90:47-116, restated per face.  The
// sequential scan carries only values that are functions of a 3-cell window
// and of the two faces of a cell, so each face can be limited independently:
//   cell bounds  qmax/qmin(c): window {c-1,c,c+1} of q1 and q0, except the
//                first cell {0,1} (:50-55) and the last {q1(n-2),q1(n-1),
//                q0(n-2)} (:62-68)
//   cell fluxes  fin/fout(c) from the unlimited donor fluxes f of its faces;
//                0 at both ends except for w, where the bottom cell sees only
//                its top face and the top cell |f| (:72-100)
// A line has n cells; face t joins cells t and t+1; P(c) is the address of
// cell c, F(t) that of face t.
// ---------------------------------------------------------------------------
struct Line {
    size_t base;    // address of cell 0
    size_t stride;  // distance between cells
    size_t fbase;   // address of face 0
    int n;          // number of cells
};

__device__ __forceinline__ float line_f(const Line &L, const float *q1, const float *U2, int t)
{
    return flux1(q1[L.base + (size_t)t * L.stride], q1[L.base + (size_t)(t + 1) * L.stride],
                 U2[L.fbase + (size_t)t * L.stride]);
}

__device__ __forceinline__ void cell_bounds(const Line &L, const float *q1, const float *q0, int c,
                                            float &qmax, float &qmin)
{
    const size_t s = L.stride;
    const size_t p = L.base + (size_t)c * s;
    if (c == 0) {
        qmax = fmaxf(fmaxf(q1[p], q1[p + s]), fmaxf(q0[p], q0[p + s]));
        qmin = fminf(fminf(q1[p], q1[p + s]), fminf(q0[p], q0[p + s]));
    } else if (c == L.n - 1) {
        qmax = fmaxf(fmaxf(q1[p - s], q1[p]), q0[p - s]);
        qmin = fminf(fminf(q1[p - s], q1[p]), q0[p - s]);
    } else {
        qmax = fmaxf(fmaxf(fmaxf(q1[p - s], q1[p]), fmaxf(q1[p + s], q0[p - s])),
                     fmaxf(q0[p], q0[p + s]));
        qmin = fminf(fminf(fminf(q1[p - s], q1[p]), fminf(q1[p + s], q0[p - s])),
                     fminf(q0[p], q0[p + s]));
    }
}

// fl = flux on the face below/left of cell c (face c-1), fr = face c.
__device__ __forceinline__ void cell_fluxes(int c, int n, bool is_w, float fl, float fr, float &fin,
                                            float &fout)
{
    if (c == 0) {
        if (is_w) {
            fin = 0.0f - fminf(0.0f, fr);
            fout = fmaxf(0.0f, fr);
        } else {
            fin = 0.0f;
            fout = 0.0f;
        }
    } else if (c == n - 1) {
        if (is_w) {
            fin = fmaxf(0.0f, fl) - fminf(0.0f, fl);
            fout = fmaxf(0.0f, fl) - fminf(0.0f, fl);
        } else {
            fin = 0.0f;
            fout = 0.0f;
        }
    } else {
        fin = fmaxf(0.0f, fl) - fminf(0.0f, fr);
        fout = fmaxf(0.0f, fr) - fminf(0.0f, fl);
    }
}

__device__ float fct_face(const Line &L, const float *q1, const float *q0, const float *U2, int t,
                          bool is_w)
{
    const float u = U2[L.fbase + (size_t)t * L.stride];
    if (!(u > 0.0f) && !(u < 0.0f)) return u;
    const int n = L.n;
    const float ft = line_f(L, q1, U2, t);
    const float fm = (t >= 1) ? line_f(L, q1, U2, t - 1) : 0.0f;
    const float fp = (t + 1 <= n - 2) ? line_f(L, q1, U2, t + 1) : 0.0f;
    float qmax_l, qmin_l, qmax_r, qmin_r, fin_l, fout_l, fin_r, fout_r;
    cell_bounds(L, q1, q0, t, qmax_l, qmin_l);
    cell_bounds(L, q1, q0, t + 1, qmax_r, qmin_r);
    cell_fluxes(t, n, is_w, fm, ft, fin_l, fout_l);
    cell_fluxes(t + 1, n, is_w, ft, fp, fin_r, fout_r);
    const float q1l = q1[L.base + (size_t)t * L.stride];
    const float q1r = q1[L.base + (size_t)(t + 1) * L.stride];
    const float eps = 1e-15f;
    if (u > 0.0f) {  // :103-107
        const float beta_out_i = (q1l - qmin_l) / (fout_l + eps);
        const float beta_in_i2 = (qmax_r - q1r) / (fin_r + eps);
        return fminf(fminf(1.0f, beta_in_i2), beta_out_i) * u;
    } else {  // :110-114
        const float beta_in_i = (qmax_l - q1l) / (fin_l + eps);
        const float beta_out_i2 = (q1r - qmin_r) / (fout_r + eps);
        return fminf(fminf(1.0f, beta_in_i), beta_out_i2) * u;
    }
}

// flux_limiter (adv_mpdata.f90:254-350): x rows for j in (jms,jme), z columns
// for i in (ims,ime) and j in (jms,jme) with w(kme) = 0, y rows everywhere.
// Faces outside those ranges keep the unlimited value.
__global__ void __launch_bounds__(BX *BY)
    k_fct(Dims d, const float *__restrict__ q0, const float *__restrict__ q1,
          const float *__restrict__ u2, const float *__restrict__ v2, const float *__restrict__ w2,
          float *__restrict__ u2l, float *__restrict__ v2l, float *__restrict__ w2l)
{
    int i, k, j;
    if (!cell_coords(d, i, k, j)) return;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const size_t c = ix(d, i, k, j);
    const size_t sk = (size_t)nx, sj = (size_t)nx * nz;
    // x face on the left of cell i = face t = i-1 of the row (:283-300)
    if (i >= 1 && j >= 1 && j <= ny - 2) {
        Line L{ix(d, 0, k, j), 1, ix(d, 1, k, j), nx};
        u2l[c] = fct_face(L, q1, q0, u2, i - 1, false);
    } else {
        u2l[c] = u2[c];
    }
    // y face on the south of row j = face t = j-1 of the column (:324-346)
    if (j >= 1) {
        Line L{ix(d, i, k, 0), sj, ix(d, i, k, 1), ny};
        v2l[c] = fct_face(L, q1, q0, v2, j - 1, false);
    } else {
        v2l[c] = v2[c];
    }
    // z face on top of level k = face t = k (:302-319)
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
        if (k <= nz - 2) {
            Line L{ix(d, i, 0, j), sk, ix(d, i, 0, j), nz};
            w2l[c] = fct_face(L, q1, q0, w2, k, true);
        } else {
            w2l[c] = 0.0f;  // :318
        }
    } else {
        w2l[c] = w2[c];
    }
}

// Ring boundary conditions (see icar_adv_apply_bc).
__global__ void k_bc_x(Dims d, float *q, int bc_x)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= d.nz || j >= d.ny) return;
    if (bc_x == 1) {
        q[ix(d, d.nx - 1, k, j)] = q[ix(d, d.nx - 2, k, j)];
    } else if (bc_x == 2) {
        q[ix(d, 0, k, j)] = q[ix(d, d.nx - 2, k, j)];
        q[ix(d, d.nx - 1, k, j)] = q[ix(d, 1, k, j)];
    }
}

__global__ void k_bc_y(Dims d, float *q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= d.nx || k >= d.nz) return;
    q[ix(d, i, k, 0)] = q[ix(d, i, k, d.ny - 2)];
    q[ix(d, i, k, d.ny - 1)] = q[ix(d, i, k, 1)];
}

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// Interior cells in a fixed order: n = i' + (nx-2)*(k + nz*j'), i' = i-1, j' = j-1.
__device__ __forceinline__ size_t interior_cell(const Dims &d, size_t n)
{
    const size_t nxi = (size_t)(d.nx - 2);
    const size_t i = n % nxi + 1;
    const size_t r = n / nxi;
    const size_t k = r % (size_t)d.nz;
    const size_t j = r / (size_t)d.nz + 1;
    return i + (size_t)d.nx * (k + (size_t)d.nz * j);
}

__global__ void __launch_bounds__(RED_THREADS)
    k_mass(Dims d, const float *__restrict__ q, const float *__restrict__ jaco,
           const float *__restrict__ dens, const float *__restrict__ dz, double *partial)
{
    const size_t total = (size_t)(d.nx - 2) * d.nz * (d.ny - 2);
    double acc = 0.0;
    for (size_t n = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; n < total;
         n += (size_t)gridDim.x * RED_THREADS) {
        const size_t c = interior_cell(d, n);
        acc += (double)q[c] * (double)jaco[c] * (double)dens[c] * (double)dz[c];
    }
    __shared__ double sh[RED_THREADS / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < RED_THREADS / 64; w++) s += sh[w];
        partial[blockIdx.x] = s;
    }
}

// Maxima of |U| over both x faces, |V| over both y faces and |W/dz| of the
// interior cells.  U holds unx faces per row (nx for the MPDATA Courant
// numbers, cell c's west face at c; nx + 1 for the standard scheme's U_m);
// dz is null when W is already divided by it.
__global__ void __launch_bounds__(RED_THREADS)
    k_courant_max(Dims d, const float *__restrict__ U, int unx, const float *__restrict__ V,
                  const float *__restrict__ W, const float *__restrict__ dz, double *partial)
{
    const size_t total = (size_t)(d.nx - 2) * d.nz * (d.ny - 2);
    double mu = 0.0, mv = 0.0, mw = 0.0;
    const size_t sj = (size_t)d.nx * d.nz;
    const size_t extra = (size_t)(unx - d.nx);
    for (size_t n = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; n < total;
         n += (size_t)gridDim.x * RED_THREADS) {
        const size_t c = interior_cell(d, n);
        const size_t fu = c + extra * (c / (size_t)d.nx);  // + one face per row before c
        mu = fmax(mu, fmax(fabs((double)U[fu]), fabs((double)U[fu + 1])));
        mv = fmax(mv, fmax(fabs((double)V[c]), fabs((double)V[c + sj])));
        mw = fmax(mw, fabs((double)(dz ? W[c] / dz[c] : W[c])));
    }
    __shared__ double sh[3][RED_THREADS / 64];
    mu = wave_max(mu);
    mv = wave_max(mv);
    mw = wave_max(mw);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = mu;
        sh[1][threadIdx.x >> 6] = mv;
        sh[2][threadIdx.x >> 6] = mw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, e = 0;
        for (int w = 0; w < RED_THREADS / 64; w++) {
            a = fmax(a, sh[0][w]);
            b = fmax(b, sh[1][w]);
            e = fmax(e, sh[2][w]);
        }
        partial[3 * blockIdx.x + 0] = a;
        partial[3 * blockIdx.x + 1] = b;
        partial[3 * blockIdx.x + 2] = e;
    }
}

// test_divergence (adv_mpdata.f90:445-475) in FP32 like the reference;
// partial[2b] = number of cells above 1e-3, partial[2b+1] = max |div|.
__global__ void __launch_bounds__(RED_THREADS)
    k_divergence(Dims d, const float *__restrict__ U, const float *__restrict__ V,
                 const float *__restrict__ W, const float *__restrict__ dz, double *partial)
{
    const size_t total = (size_t)(d.nx - 2) * d.nz * (d.ny - 2);
    const size_t sj = (size_t)d.nx * d.nz;
    double cnt = 0.0, mx = 0.0;
    for (size_t n = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; n < total;
         n += (size_t)gridDim.x * RED_THREADS) {
        const size_t c = interior_cell(d, n);
        const int k = (int)((c / (size_t)d.nx) % (size_t)d.nz);
        const float du = U[c + 1] - U[c];
        const float dv = V[c + sj] - V[c];
        const float dw = (k == 0) ? W[c] / dz[c] : (W[c] - W[c - d.nx]) / dz[c];
        const float a = fabsf(du + dv + dw);
        if (a > 1e-3f) cnt += 1.0;
        mx = fmax(mx, (double)a);
    }
    __shared__ double sh[2][RED_THREADS / 64];
    cnt = wave_sum(cnt);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = cnt;
        sh[1][threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < RED_THREADS / 64; w++) {
            a += sh[0][w];
            b = fmax(b, sh[1][w]);
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

__global__ void k_copy(size_t n, const float4 *__restrict__ src, float4 *__restrict__ dst,
                       const float *__restrict__ s1, float *__restrict__ d1, size_t n4)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t p = t; p < n4; p += stride) dst[p] = src[p];
    for (size_t p = n4 * 4 + t; p < n; p += stride) d1[p] = s1[p];
}

int red_blocks(const Dims &d)
{
    const size_t total = (size_t)(d.nx - 2) * d.nz * (d.ny - 2);
    size_t b = (total + RED_THREADS - 1) / RED_THREADS;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    if (b < 1) b = 1;
    return (int)b;
}

}  // namespace

void launch_courant(hipStream_t s, Dims d, float dx, float dt, int advect_density,
                    const DomainFields &f, const StepFields &o)
{
    hipLaunchKernelGGL(k_courant, cell_grid(d), dim3(BX, BY), 0, s, d, dx, dt, advect_density, f, o);
}

void launch_upwind(hipStream_t s, Dims d, const float *qin, const float *U, const float *V,
                   const float *W, const float *G, const float *Gz, float *q)
{
    hipLaunchKernelGGL(k_upwind, cell_grid(d), dim3(BX, BY), 0, s, d, qin, U, V, W, G, Gz, q);
}

void launch_pseudo(hipStream_t s, Dims d, const float *q1, const StepFields &sf, const float *dz,
                   float *u2, float *v2, float *w2)
{
    hipLaunchKernelGGL(k_pseudo, cell_grid(d), dim3(BX, BY), 0, s, d, q1, sf.U, sf.V, sf.Wd, sf.G,
                       dz, u2, v2, w2);
}

void launch_fct(hipStream_t s, Dims d, const float *q0, const float *q1, const float *u2,
                const float *v2, const float *w2, float *u2l, float *v2l, float *w2l)
{
    hipLaunchKernelGGL(k_fct, cell_grid(d), dim3(BX, BY), 0, s, d, q0, q1, u2, v2, w2, u2l, v2l,
                       w2l);
}

void launch_bc(hipStream_t s, Dims d, float *q, int bc_x, int bc_y)
{
    if (bc_x == 1 || bc_x == 2)
        hipLaunchKernelGGL(k_bc_x, dim3((d.nz + 63) / 64, d.ny), dim3(64), 0, s, d, q, bc_x);
    if (bc_y == 2)
        hipLaunchKernelGGL(k_bc_y, dim3((d.nx + 255) / 256, d.nz), dim3(256), 0, s, d, q);
}

int mass_partials_blocks(Dims d) { return red_blocks(d); }

void launch_mass(hipStream_t s, Dims d, const float *q, const float *jaco, const float *density,
                 const float *dz, double *partial)
{
    hipLaunchKernelGGL(k_mass, dim3(red_blocks(d)), dim3(RED_THREADS), 0, s, d, q, jaco, density,
                       dz, partial);
}

void launch_courant_max(hipStream_t s, Dims d, const float *U, int unx, const float *V,
                        const float *W, const float *dz, double *partial)
{
    hipLaunchKernelGGL(k_courant_max, dim3(red_blocks(d)), dim3(RED_THREADS), 0, s, d, U, unx, V,
                       W, dz, partial);
}

void launch_divergence(hipStream_t s, Dims d, const float *U, const float *V, const float *W,
                       const float *dz, double *partial)
{
    hipLaunchKernelGGL(k_divergence, dim3(red_blocks(d)), dim3(RED_THREADS), 0, s, d, U, V, W, dz,
                       partial);
}

void launch_copy(hipStream_t s, size_t n, const float *src, float *dst)
{
    const size_t n4 = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15)
                          ? 0
                          : n / 4;
    size_t blocks = (n / 4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, n,
                       reinterpret_cast<const float4 *>(src), reinterpret_cast<float4 *>(dst), src,
                       dst, n4);
}

}  // namespace icaradv
