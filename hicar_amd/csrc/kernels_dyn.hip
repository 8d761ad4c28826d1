// The steps ICAR runs next to advect, on the device copies of the fields the
// advection handle already holds:
//   k_balance_w   balance_uvw (src/physics/wind.f90:88-164, update=.false.):
//                 the horizontal mass-flux divergence (calc_divergence with
//                 horz_only, :201-234) and w integrated up each column so
//                 that it balances (calc_w, :166-199).  One thread per
//                 column, threads along i (coalesced rows), the level below
//                 carried in registers.
//   k_dt_partials compute_dt (src/main/time_step.f90:51-168) with
//                 use_density=.false.: per block the maxima of |u|, |v|, |w|
//                 over the whole arrays and of the tile's per-cell Courant
//                 rate; the host takes the maximum over blocks and finishes
//                 in FP32 as the reference does (a maximum is exact in any
//                 order, so the result is the reference's bits).
//   k_forcing     the boundary-forced branch of apply_forcing
//                 (src/objects/domain_obj.f90:3493-3503) for one species.
//   k_strip       the halo strips of the species for the halo exchange.
//   k_clamp       enforce_limits (domain_obj.f90:3092-3150) for one species.
// Every expression keeps the Fortran's evaluation order; this unit is
// compiled with -ffp-contract=off.
#include "../../include/icar_adv.h"
#include "common.hpp"

namespace icaradv {

namespace {

// u_met of x face fi (0..nx) of row (k, j), wind.f90:222-228
__device__ __forceinline__ float u_met(const Dims &d, int fi, int k, int j, int adv_den,
                                       const float *__restrict__ u, const float *__restrict__ ju,
                                       const float *__restrict__ rho)
{
    const size_t f = (size_t)fi + (size_t)(d.nx + 1) * ((size_t)k + (size_t)d.nz * j);
    if (!adv_den) return u[f] * ju[f];
    const size_t c = ix(d, 0, k, j);
    if (fi == 0) return u[f] * ju[f] * (1.5f * rho[c] - 0.5f * rho[c + 1]);
    if (fi == d.nx) return u[f] * ju[f] * (1.5f * rho[c + d.nx - 1] - 0.5f * rho[c + d.nx - 2]);
    return u[f] * ju[f] * (rho[c + fi - 1] + rho[c + fi]) / 2.0f;
}

// v_met of y face fj (0..ny) of column (i, k), wind.f90:225-227
__device__ __forceinline__ float v_met(const Dims &d, int i, int k, int fj, int adv_den,
                                       const float *__restrict__ v, const float *__restrict__ jv,
                                       const float *__restrict__ rho)
{
    const size_t f = ix(d, i, k, fj);
    if (!adv_den) return v[f] * jv[f];
    const size_t sj = (size_t)d.nx * d.nz;
    const size_t c0 = ix(d, i, k, 0);
    if (fj == 0) return v[f] * jv[f] * (1.5f * rho[c0] - 0.5f * rho[c0 + sj]);
    if (fj == d.ny)
        return v[f] * jv[f] * (1.5f * rho[c0 + (d.ny - 1) * sj] - 0.5f * rho[c0 + (d.ny - 2) * sj]);
    return v[f] * jv[f] * (rho[f - sj] + rho[f]) / 2.0f;
}

__global__ void __launch_bounds__(256)
    k_balance_w(Dims d, int i_s, int i_e, int j_s, int j_e, float dx, int adv_den, DomainFields f,
                float *__restrict__ w)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= d.nx || j >= d.ny) return;
    const size_t sk = (size_t)d.nx;
    const size_t c0 = ix(d, i, 0, j);
    if (i < i_s || i > i_e || j < j_s || j > j_e) {  // w = 0 outside the balanced columns (:179)
        for (int k = 0; k < d.nz; k++) w[c0 + sk * k] = 0.0f;
        return;
    }
    const float *__restrict__ rho = f.density;
    const float *__restrict__ dz = f.dz;
    const float *__restrict__ jw = f.jaco_w;
    float wb = 0.0f, rib = 0.0f, jwb = 0.0f;  // w, rho_i and jacobian_w of the level below
    for (int k = 0; k < d.nz; k++) {
        const size_t c = c0 + sk * k;
        const float dU = u_met(d, i + 1, k, j, adv_den, f.u, f.jaco_u, rho) -
                         u_met(d, i, k, j, adv_den, f.u, f.jaco_u, rho);
        const float dV = v_met(d, i, k, j + 1, adv_den, f.v, f.jaco_v, rho) -
                         v_met(d, i, k, j, adv_den, f.v, f.jaco_v, rho);
        const float div = (dU + dV) / dx;  // :232
        float wc;
        if (adv_den) {
            // rho_i(k) = (rho(k)dz(k) + rho(k+1)dz(k+1)) / (dz(k)+dz(k+1)), :177
            const float ri = (k < d.nz - 1)
                                 ? (rho[c] * dz[c] + rho[c + sk] * dz[c + sk]) / (dz[c] + dz[c + sk])
                                 : 0.0f;
            if (k == 0)
                wc = 0.0f - div * dz[c] / (jw[c] * ri);                           // :183-184
            else if (k == d.nz - 1)
                wc = ((wb * rib * jwb) - div * dz[c]) / (jw[c] * rho[c]);          // :185-188
            else
                wc = ((wb * rib * jwb) - div * dz[c]) / (jw[c] * ri);              // :189-192
            rib = ri;
        } else {
            if (k == 0)
                wc = (0.0f - div * dz[c]) / (jw[c]);                               // :195
            else
                wc = (wb * jwb - div * dz[c]) / (jw[c]);                           // :197-198
        }
        w[c] = wc;
        wb = wc;
        jwb = jw[c];
    }
}

constexpr int DT_BLOCKS = 1024, DT_THREADS = 256;

__device__ __forceinline__ float block_max(float v, float *sm)
{
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = DT_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    const float r = sm[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(DT_THREADS)
    k_dt_partials(Dims d, int its, int ite, int jts, int jte, float dx, const float *__restrict__ u,
                  const float *__restrict__ v, const float *__restrict__ w,
                  const float *__restrict__ dzl, float *__restrict__ part)
{
    __shared__ float sm[DT_THREADS];
    const size_t stride = (size_t)gridDim.x * DT_THREADS;
    const size_t t0 = (size_t)blockIdx.x * DT_THREADS + threadIdx.x;
    const size_t nu = (size_t)(d.nx + 1) * d.nz * d.ny, nv = (size_t)d.nx * d.nz * (d.ny + 1);
    const size_t nw = (size_t)d.nx * d.nz * d.ny;
    float mu = 0.0f, mv = 0.0f, mw = 0.0f, m3 = 0.0f;
    for (size_t p = t0; p < nu; p += stride) mu = fmaxf(mu, fabsf(u[p]));
    for (size_t p = t0; p < nv; p += stride) mv = fmaxf(mv, fabsf(v[p]));
    for (size_t p = t0; p < nw; p += stride) mw = fmaxf(mw, fabsf(w[p]));
    // the tile, :101-126: i fastest so that neighbouring threads read a row
    const int ti = ite - its + 1, tj = jte - jts + 1;
    const size_t nt = (size_t)ti * d.nz * tj;
    for (size_t p = t0; p < nt; p += stride) {
        const int i = its + (int)(p % (size_t)ti);
        const size_t r = p / (size_t)ti;
        const int k = (int)(r % (size_t)d.nz);
        const int j = jts + (int)(r / (size_t)d.nz);
        const size_t fu = (size_t)i + (size_t)(d.nx + 1) * ((size_t)k + (size_t)d.nz * j);
        const size_t c = ix(d, i, k, j);
        const size_t cb = k == 0 ? c : c - (size_t)d.nx;  // zoffset, :104-108
        const float a = fmaxf(fabsf(u[fu]), fabsf(u[fu + 1])) / dx;
        const float b = fmaxf(fabsf(v[c]), fabsf(v[c + (size_t)d.nx * d.nz])) / dx;
        const float e = fmaxf(fabsf(w[c]), fabsf(w[cb])) / dzl[k];
        m3 = fmaxf(m3, fmaxf(fmaxf(a, b), e));
    }
    mu = block_max(mu, sm);
    mv = block_max(mv, sm);
    mw = block_max(mw, sm);
    m3 = block_max(m3, sm);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = mu;
        part[4 * blockIdx.x + 1] = mv;
        part[4 * blockIdx.x + 2] = mw;
        part[4 * blockIdx.x + 3] = m3;
    }
}

}  // namespace

void launch_balance_w(hipStream_t s, Dims d, int i_s, int i_e, int j_s, int j_e, float dx,
                      int adv_den, const DomainFields &f, float *w)
{
    hipLaunchKernelGGL(k_balance_w, dim3((d.nx + 63) / 64, (d.ny + 3) / 4), dim3(64, 4), 0, s, d,
                       i_s, i_e, j_s, j_e, dx, adv_den, f, w);
}

int dt_partial_blocks() { return DT_BLOCKS; }

void launch_dt_partials(hipStream_t s, Dims d, int its, int ite, int jts, int jte, float dx,
                        const float *u, const float *v, const float *w, const float *dzl,
                        float *part)
{
    hipLaunchKernelGGL(k_dt_partials, dim3(DT_BLOCKS), dim3(DT_THREADS), 0, s, d, its, ite, jts,
                       jte, dx, u, v, w, dzl, part);
}

}  // namespace icaradv

namespace icaradv {

namespace {

// apply_forcing for a boundary-forced 3-D variable (domain_obj.f90:3493-3503):
// the forcing advances by its tendency, then the value takes it where the
// filter is 1 and relaxes towards it where the filter is below 1.  dt is
// double precision in the reference (dt%seconds()), so the tendency and the
// relaxation term are formed in FP64 and rounded to FP32 on assignment, as
// the Fortran does; forcing - value is an FP32 difference.
__global__ void __launch_bounds__(256)
    k_forcing(Dims d, double dt, const float *__restrict__ filt, float *__restrict__ q,
              float *__restrict__ fhi, const float *__restrict__ dqdt)
{
    const size_t n = (size_t)d.nx * d.nz * d.ny;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const float f = (float)((double)fhi[c] + ((double)dqdt[c] * dt));
        fhi[c] = f;
        const size_t i = c % (size_t)d.nx, j = c / ((size_t)d.nx * d.nz);
        const float rf = filt[j * (size_t)d.nx + i];
        if (rf == 1.0f)
            q[c] = f;
        else if (rf < 1.0f)
            q[c] = (float)((double)q[c] + ((double)rf * dt / 3600.0) * (double)(f - q[c]));
    }
}

}  // namespace

void launch_forcing(hipStream_t s, Dims d, double dt, const float *filt, float *q, float *fhi,
                    const float *dqdt)
{
    const size_t n = (size_t)d.nx * d.nz * d.ny;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_forcing, dim3((unsigned)blocks), dim3(256), 0, s, d, dt, filt, q, fhi,
                       dqdt);
}

}  // namespace icaradv

namespace icaradv {

namespace {

// Copies one rectangular strip [i0, i0+ni) x [j0, j0+nj), all levels, of
// every species between the species arrays and a packed buffer, or from one
// handle's species to another's (both strips the same shape).  Elements are
// moved as raw 4- or 2-byte words, so 16-bit storage travels as it is held.
// Buffer order: species, then j, k, i (i fastest): each species' strip is a
// contiguous sub-box of its (i, k, j) array.
__global__ void __launch_bounds__(256)
    k_strip(HaloStrips hs, int nz, long long per_species_cells, int mode)
{
    const long long total = per_species_cells * hs.n;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < total;
         t += (long long)gridDim.x * 256) {
        const int s = (int)(t / per_species_cells);
        long long r = t - (long long)s * per_species_cells;
        const int ii = (int)(r % hs.ni);
        r /= hs.ni;
        const int k = (int)(r % nz);
        const int jj = (int)(r / nz);
        const size_t a = (size_t)(hs.a_i0 + ii) + (size_t)hs.a_nx * ((size_t)k + (size_t)nz * (hs.a_j0 + jj));
        const size_t b = (size_t)(hs.b_i0 + ii) + (size_t)hs.b_nx * ((size_t)k + (size_t)nz * (hs.b_j0 + jj));
        const int es = hs.esize[s];
        char *boff = (char *)hs.buf + (size_t)hs.buf_off[s];
        if (mode == 0) {  // species a -> buffer
            if (es == 4) ((float *)boff)[t - (long long)s * per_species_cells] = ((const float *)hs.a[s])[a];
            else ((uint16_t *)boff)[t - (long long)s * per_species_cells] = ((const uint16_t *)hs.a[s])[a];
        } else if (mode == 1) {  // buffer -> species b
            if (es == 4) ((float *)hs.b[s])[b] = ((const float *)boff)[t - (long long)s * per_species_cells];
            else ((uint16_t *)hs.b[s])[b] = ((const uint16_t *)boff)[t - (long long)s * per_species_cells];
        } else {  // species a -> species b
            if (es == 4) ((float *)hs.b[s])[b] = ((const float *)hs.a[s])[a];
            else ((uint16_t *)hs.b[s])[b] = ((const uint16_t *)hs.a[s])[a];
        }
    }
}

}  // namespace

void launch_strip(hipStream_t s, const HaloStrips &hs, int nz, int mode)
{
    const long long per = (long long)hs.ni * hs.nj * nz;
    const long long total = per * hs.n;
    if (total <= 0) return;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_strip, dim3((unsigned)blocks), dim3(256), 0, s, hs, nz, per, mode);
}

}  // namespace icaradv

namespace icaradv {

namespace {

// enforce_limits (domain_obj.f90:3092-3150, the data_3d branch): where a
// species is below zero it becomes zero.  -0.0 and NaN are not below zero
// and stay.  16-bit storage is judged on its bits: a set sign bit with a
// non-zero magnitude that is not a NaN.
__global__ void __launch_bounds__(256)
    k_clamp(size_t n, void *__restrict__ q, int storage)
{
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        if (storage == ICAR_ADV_STORE_F32) {
            float *f = static_cast<float *>(q);
            const float v = f[p];
            if (v < 0.0f) f[p] = 0.0f;
        } else {
            uint16_t *h = static_cast<uint16_t *>(q);
            const uint16_t v = h[p];
            const bool bf = storage == ICAR_ADV_STORE_BF16 || storage == ICAR_ADV_STORE_BF16_SR;
            const uint16_t expm = bf ? 0x7f80u : 0x7c00u, manm = bf ? 0x007fu : 0x03ffu;
            const bool nan = (v & expm) == expm && (v & manm) != 0;
            if ((v & 0x8000u) && (v & 0x7fffu) && !nan) h[p] = 0;
        }
    }
}

}  // namespace

void launch_clamp(hipStream_t s, size_t n, void *q, int storage)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_clamp, dim3((unsigned)blocks), dim3(256), 0, s, n, q, storage);
}

}  // namespace icaradv
