// The standard advection scheme (&physics adv=1) on the device:
//   k_std_wind    adv_std_compute_wind (src/physics/advect.f90:444-492): the
//                 mass fluxes U_m, V_m, W_m and the divisors jacobian*rho and
//                 dz*jacobian*rho of the flux-form update.
//   k_std_stage   one adv_std_advect3d call without the limiter (advect.f90:
//                 259-320): flux3 of the stage state with the winds times
//                 t_factor, applied to the base state.  Mode 1 also writes
//                 the first upwind half step of upwind_flux3 (dumb_q,
//                 adv_fluxcorr.f90:172-175), which depends on the base state
//                 only, so RK3 gets it from its first stage's loads; mode 2
//                 writes only that.
//   k_std_scale   WRF_flux_corr (adv_fluxcorr.f90:37-127) up to scale_in and
//                 scale_out, on the tile plus one ring of cells.
//   k_std_final   the normalised, re-merged fluxes (:129-150) and the update.
//   k_std_ring    the cells outside the tile, which every call leaves at the
//                 base state (qfluxes = qold, :298).
// The stage, scale and final kernels march up columns (see Lane below) and
// loop over a batch of species; each x and z face flux is computed once and
// shared, the y faces by both rows.  Every evaluation is the same IEEE FP32
// expression in the Fortran's order (this unit is compiled with
// -ffp-contract=off), so the results are the bits of a strict build of the
// reference.  Fortran MAX/MIN of two reals are the selects a > b ? a : b and
// a < b ? a : b.
#include <algorithm>

#include "common.hpp"

namespace icaradv {

namespace {

// Element offsets are 32-bit and byte offsets are formed in 32 bits (the
// host limits a domain to 2^29 cells), so that a load is a uniform base in
// scalar registers plus one 32-bit vector offset.
typedef uint32_t off;
typedef ptrdiff_t soff;  // signed offsets where loads go through 64-bit pointers

template <class T>
struct G {
    const T *p;
    __device__ __forceinline__ T operator[](off i) const
    {
        typedef __attribute__((address_space(1))) const char gcchar;
        typedef __attribute__((address_space(1))) const T gct;
        return *(gct *)((gcchar *)p + (uint32_t)(i * (uint32_t)sizeof(T)));
    }
};

// (scale_in, scale_out) pairs as two adjacent floats
template <>
struct G<float2> {
    const float2 *p;
    __device__ __forceinline__ float2 operator[](off i) const
    {
        const G<float> f{reinterpret_cast<const float *>(p)};
        return make_float2(f[2 * i], f[2 * i + 1]);
    }
};

__device__ __forceinline__ void put(float *p, off i, float v)
{
    typedef __attribute__((address_space(1))) char gchar;
    typedef __attribute__((address_space(1))) float gt;
    *(gt *)((gchar *)p + (uint32_t)(i * 4u)) = v;
}

__device__ __forceinline__ void put(float2 *p, off i, float2 v)
{
    float *f = reinterpret_cast<float *>(p);
    put(f, 2 * i, v.x);
    put(f, 2 * i + 1, v.y);
}

// A block of cells [i_lo, i_hi] x [j_lo, j_hi] (0-based), all levels, cut
// into chunks of kch levels (one per grid z index).
struct Region {
    int i_lo, i_hi, j_lo, j_hi, kch;
    int nbx, nby, nbz;  // blocks along i, j and k (set by with_blocks)
    int sg;             // species per block (a group of the batch)
};

__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }

// flux3's face formulas (advect.f90:150-178): m3..p2 are q at offsets -3..+2
// from the cell on the positive side of the face.
__device__ __forceinline__ float face_up(float u, float m1, float p0)
{
    return ((u + fabsf(u)) * m1 + (u - fabsf(u)) * p0) / 2.0f;
}

__device__ __forceinline__ float face_o3(float u, float m2, float m1, float p0, float p1)
{
    float f = 7.0f * (p0 + m1) - (p1 + m2);
    f = (u * f) / 12.0f;
    f = f - (fabsf(u) / 12.0f) * (3.0f * (p0 - m1) - (p1 - m2));
    return f;
}

__device__ __forceinline__ float face_o5(float u, float m3, float m2, float m1, float p0,
                                         float p1, float p2)
{
    float f = 37.0f * (p0 + m1) - 8.0f * (p1 + m2) + (p2 + m3);
    f = (u * f) / 60.0f;
    f = f - (fabsf(u) / 60.0f) * (10.0f * (p0 - m1) - 5.0f * (p1 - m2) + (p2 - m3));
    return f;
}

// horizontal face whose positive-side cell is c, stride s (q: a raw pointer
// with signed offsets, or a G<float> with 32-bit ones)
template <int H, class A, class I, class J>
__device__ __forceinline__ float face_h(float u, const A &q, I c, J stride)
{
    const I s = (I)stride;
    if constexpr (H == 1) {
        return face_up(u, q[c - s], q[c]);
    } else if constexpr (H == 3) {
        return face_o3(u, q[c - 2 * s], q[c - s], q[c], q[c + s]);
    } else {
        return face_o5(u, q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s]);
    }
}

// one upwind half step of upwind_flux3 (adv_fluxcorr.f90:160-165)
__device__ __forceinline__ float half_up(float u, float m1, float p0)
{
    return 0.5f * ((u + fabsf(u)) * m1 + (u - fabsf(u)) * p0) / 2.0f;
}

// vertical face kf (1..nz-1) from the window z[m] = q(kf-3+m), m = 0..5
// (advect.f90:180-241): the faces next to the ground and the top stay
// upwind, 5th order falls back to 3rd one face further in.  Window entries
// outside the column are never read by the formula chosen.
template <int V>
__device__ __forceinline__ float face_vw(float w, const float (&z)[6], int kf, int nz)
{
    if (V == 1 || kf == 1 || kf == nz - 1) return face_up(w, z[2], z[3]);
    if (V == 5 && kf >= 3 && kf <= nz - 3) return face_o5(w, z[0], z[1], z[2], z[3], z[4], z[5]);
    return face_o3(w, z[1], z[2], z[3], z[4]);
}

// The flux-form update of adv_std_advect3d (advect.f90:301-309).
__device__ __forceinline__ float flux_update(float q0, float xw, float xe, float ys, float yn,
                                             float zb, float zt, float rj, float rdj)
{
    float r = q0 - ((xe - xw) + (yn - ys)) / rj;
    r = r - (zt - zb) / rdj;
    return r;
}

// A merged flux m through a face, normalised as adv_fluxcorr.f90:129-145
// does: pos / neg are the (scale_in, scale_out) of the cells on the positive
// and negative side of the face.
__device__ __forceinline__ float limit_face(float m, float2 pos, float2 neg)
{
    if (m > 0.0f) return fmn(pos.x, neg.y) * m;
    if (m < 0.0f) return fmn(pos.y, neg.x) * m;
    return m;
}

// Both half steps of upwind_flux3 through one face (adv_fluxcorr.f90:
// 158-185): the first from q, the second added from dumb_q.
__device__ __forceinline__ float up2(float u, float qm, float qp, float dm, float dp)
{
    float f = half_up(u, qm, qp);
    f = f + half_up(u, dm, dp);
    return f;
}

// The kernels march up the columns of a region, a chunk of levels per
// thread: a wave holds 64 columns along i and updates the first 63; each
// lane computes the fluxes through the x face on the west side of its column
// and takes the east face from the next lane (lane 63 only supplies that
// face).  The face under level k is the one computed above level k-1,
// carried in registers, so each z face is computed once too, apart from the
// face under the first level of a chunk; the y faces are computed by both
// rows.
constexpr int CW = 63;  // columns a wave updates
constexpr int RB = 4;   // rows (waves) per block

struct Lane {
    int i, j, fi;    // cell, and the column whose west face the lane supplies
    int k0, k1;      // the chunk of levels [k0, k1)
    int s0, s1;      // the species [s0, s1) of the batch
    bool cell;       // the lane updates cell (i, j)
    off c0, u0;      // level 0 of column fi in cells and in the x-face array
};

// Blocks are dealt round-robin over the 8 XCDs; renumber them so that each
// XCD takes a contiguous range (neighbouring rows of tiles, whose y
// stencils overlap, then share that XCD's L2).  The launch is 1-D.
__device__ __forceinline__ void block_xyz(const Region &r, int &bx, int &by, int &bz, int &bg)
{
    const unsigned L = blockIdx.x, B = gridDim.x;
    const unsigned q = B / 8, rem = B % 8, x = L % 8, slot = L / 8;
    const unsigned id = x < rem ? x * (q + 1) + slot : rem * (q + 1) + (x - rem) * q + slot;
    const unsigned per = (unsigned)r.nbx * (unsigned)r.nby * (unsigned)r.nbz;
    const unsigned t = id % per;
    bg = (int)(id / per);  // the species group is the slowest index
    bx = (int)(t % (unsigned)r.nbx);
    by = (int)((t / (unsigned)r.nbx) % (unsigned)r.nby);
    bz = (int)(t / ((unsigned)r.nbx * (unsigned)r.nby));
}

__device__ __forceinline__ Lane lane_of(const Dims &d, const Region &r)
{
    Lane l;
    int bx, by, bz, bg;
    block_xyz(r, bx, by, bz, bg);
    l.s0 = bg * r.sg;
    l.s1 = l.s0 + r.sg;
    const int lane = threadIdx.x;
    l.i = r.i_lo + bx * CW + lane;
    l.j = r.j_lo + by * RB + (int)threadIdx.y;
    l.fi = min(l.i, r.i_hi + 1);
    l.cell = lane < CW && l.i <= r.i_hi;
    l.k0 = bz * r.kch;
    l.k1 = min(d.nz, l.k0 + r.kch);
    l.c0 = (off)l.fi + (off)d.nx * (off)d.nz * l.j;
    l.u0 = (off)l.fi + (off)(d.nx + 1) * (off)d.nz * l.j;
    return l;
}

// q(k0-3 .. k0+3) of a column (zero outside it): w[0..5] is the stencil of
// the face under level k0, w[1..6] the window the march starts with
__device__ __forceinline__ void load_window7(float (&w)[7], G<float> q, off c0, int k0, int nz,
                                             off sz)
{
#pragma unroll
    for (int m = 0; m < 7; m++) {
        const int k = k0 - 3 + m;
        w[m] = (k >= 0 && k < nz) ? q[c0 + (off)k * sz] : 0.0f;
    }
}

template <int V>
__device__ __forceinline__ float face_under(float w, const float (&w7)[7], int k0, int nz)
{
    const float z[6] = {w7[0], w7[1], w7[2], w7[3], w7[4], w7[5]};
    return face_vw<V>(w, z, k0, nz);
}

// one thread per cell: 64 along i by 4 along k per block, a row j per block column
struct Cell {
    int i, k, j;
    off c;
};

__device__ __forceinline__ Cell this_cell(const Dims &d)
{
    Cell p;
    p.i = blockIdx.x * 64 + threadIdx.x;
    p.k = blockIdx.y * 4 + threadIdx.y;
    p.j = blockIdx.z;
    p.c = (off)p.i + (off)d.nx * ((off)p.k + (off)d.nz * p.j);
    return p;
}

// ---------------------------------------------------------------------------
// adv_std_compute_wind
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_std_wind(Dims d, float dx, float dt, int adv_den, DomainFields f, float *__restrict__ Um,
               float *__restrict__ Vm, float *__restrict__ Wm, float *__restrict__ Rj,
               float *__restrict__ Rdj)
{
    const Cell p = this_cell(d);
    if (p.i >= d.nx || p.k >= d.nz) return;
    const off c = p.c, sj = (off)d.nx * d.nz, sz = d.nx;
    const float r0 = adv_den ? f.density[c] : 1.0f;  // :472-473
    const off fu = (off)p.i + (off)(d.nx + 1) * ((off)p.k + (off)d.nz * p.j);
    if (p.i >= 1) {  // :478-479, U_m(ims+1:ime)
        const float rw = adv_den ? f.density[c - 1] : 1.0f;
        Um[fu] = f.u[fu] * dt * (r0 + rw) * 0.5f * f.jaco_u[fu] / dx;
    } else {
        Um[fu] = 0.0f;
    }
    if (p.i == d.nx - 1) Um[fu + 1] = 0.0f;
    if (p.j >= 1) {  // :480-481, V_m(:,:,jms+1:jme)
        const float rs = adv_den ? f.density[c - sj] : 1.0f;
        Vm[c] = f.v[c] * dt * (r0 + rs) * 0.5f * f.jaco_v[c] / dx;
    } else {
        Vm[c] = 0.0f;
    }
    if (p.j == d.ny - 1) Vm[c + sj] = 0.0f;
    if (p.k < d.nz - 1) {  // :475-476 rho_i, :483
        const float ra = adv_den ? f.density[c + sz] : 1.0f;
        const float ri = (r0 * f.dz[c] + ra * f.dz[c + sz]) / (f.dz[c] + f.dz[c + sz]);
        Wm[c] = f.w[c] * dt * f.jaco_w[c] * ri;
    } else {  // :484
        Wm[c] = f.w[c] * dt * f.jaco_w[c] * r0;
    }
    Rj[c] = f.jaco[c] * r0;             // jaco*rho of :304
    Rdj[c] = f.dz[c] * f.jaco[c] * r0;  // dz*jaco*rho of :309
}

// ---------------------------------------------------------------------------
// one RK3 stage / Euler step without the limiter (MODE 0, tile), the same
// plus the dumb_q half step on the tile and one ring (MODE 1, that region),
// or dumb_q alone (MODE 2)
// ---------------------------------------------------------------------------
// vertical face kf (1..nz-1) whose upper cell is c, read from memory
template <int V, class A, class I>
__device__ __forceinline__ float face_vg(float w, const A &q, I c, int kf, int nz, I s)
{
    if (V == 1 || kf == 1 || kf == nz - 1) return face_up(w, q[c - s], q[c]);
    if (V == 5 && kf >= 3 && kf <= nz - 3)
        return face_o5(w, q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s]);
    return face_o3(w, q[c - 2 * s], q[c - s], q[c], q[c + s]);
}

// One thread per cell of the region (64 along i by 4 levels per block, a
// row per block column), every face computed by each cell that uses it.  A
// stage has little work per face, so the parallelism of this layout is
// worth more than the shared faces a column march would give it.
template <int H, int V, int MODE>
__global__ void __launch_bounds__(256) k_std_stage_cell(StdGeom g, float t, StdBatch b, Region r)
{
    const Dims d = g.d;
    const int i = r.i_lo + (int)blockIdx.x * 64 + (int)threadIdx.x;
    const int k = (int)blockIdx.y * 4 + (int)threadIdx.y;
    const int j = r.j_lo + (int)blockIdx.z;
    if (i > r.i_hi || k >= d.nz) return;
    const float *__restrict__ Um = g.Um, *__restrict__ Vm = g.Vm, *__restrict__ Wm = g.Wm;
    const float *__restrict__ Rj = g.Rj, *__restrict__ Rdj = g.Rdj;
    const bool tile = i >= g.its && i <= g.ite && j >= g.jts && j <= g.jte;
    const soff sj = (soff)d.nx * d.nz, sz = d.nx;
    const soff c = (soff)i + (soff)d.nx * ((soff)k + (soff)d.nz * j);
    const soff fu = (soff)i + (soff)(d.nx + 1) * ((soff)k + (soff)d.nz * j);
    const float uw = Um[fu], ue = Um[fu + 1], vs = Vm[c], vn = Vm[c + sj];
    const float wb = k > 0 ? Wm[c - sz] : 0.0f, wt = Wm[c];
    const float rj = Rj[c], rdj = Rdj[c];
    for (int s = 0; s < b.n; s++) {
        const float *__restrict__ qin = b.qin[s];
        const float *__restrict__ qo = b.qold[s];
        const float q0 = qo[c];
        if (MODE <= 1 && tile) {  // flux3 with the winds times t (advect.f90:285-287)
            const float fxw = face_h<H>(uw * t, qin, c, 1);
            const float fxe = face_h<H>(ue * t, qin, c + 1, 1);
            const float fys = face_h<H>(vs * t, qin, c, sj);
            const float fyn = face_h<H>(vn * t, qin, c + sj, sj);
            const float fzb = k == 0 ? 0.0f : face_vg<V>(wb * t, qin, c, k, d.nz, sz);  // :252
            const float fzt = k == d.nz - 1 ? qin[c] * (wt * t)                          // :253
                                            : face_vg<V>(wt * t, qin, c + sz, k + 1, d.nz, sz);
            b.out[s][c] = flux_update(q0, fxw, fxe, fys, fyn, fzb, fzt, rj, rdj);
        }
        if (MODE >= 1) {  // dumb_q (adv_fluxcorr.f90:158-175)
            const float hxw = half_up(uw, qo[c - 1], q0);
            const float hxe = half_up(ue, q0, qo[c + 1]);
            const float hys = half_up(vs, qo[c - sj], q0);
            const float hyn = half_up(vn, q0, qo[c + sj]);
            const float hzb = k == 0 ? 0.0f : half_up(wb, qo[c - sz], q0);
            const float hzt = k == d.nz - 1 ? 0.5f * q0 * wt : half_up(wt, q0, qo[c + sz]);
            b.dq[s][c] = flux_update(q0, hxw, hxe, hys, hyn, hzb, hzt, rj, rdj);
        }
    }
}

// ---------------------------------------------------------------------------
// WRF_flux_corr up to scale_in / scale_out, on the tile plus one ring
// ---------------------------------------------------------------------------
template <int H, int V>
__global__ void __launch_bounds__(64 * RB) k_std_scale(StdGeom g, StdBatch b, Region r)
{
    const Dims d = g.d;
    const Lane l = lane_of(d, r);
    if (l.j > r.j_hi) return;
    const G<float> Um{g.Um}, Vm{g.Vm}, Wm{g.Wm}, Rj{g.Rj}, Rdj{g.Rdj};
    const off sj = (off)d.nx * d.nz, sz = d.nx;
    for (int s = l.s0; s < min(l.s1, b.n); s++) {
        const G<float> qin{b.qin[s]}, qo{b.qold[s]}, dq{b.dq[s]};
        float zw[6], w7[7];
        load_window7(w7, qin, l.c0, l.k0, d.nz, sz);
#pragma unroll
        for (int m = 0; m < 6; m++) zw[m] = w7[m + 1];
        const off ck = l.c0 + (off)l.k0 * sz;
        float ob = 0.0f, o0 = qo[ck], d0 = dq[ck];  // qold below, qold and dumb_q here
        float uzb = 0.0f, hzb = 0.0f, wb = 0.0f;    // the face under the level
        if (l.k0 > 0) {
            wb = Wm[ck - sz];
            ob = qo[ck - sz];
            uzb = up2(wb, ob, o0, dq[ck - sz], d0);
            hzb = face_under<V>(wb, w7, l.k0, d.nz);
        }
#pragma unroll 1
        for (int k = l.k0; k < l.k1; k++) {
            const off c = l.c0 + (off)k * sz;
            const bool top = k == d.nz - 1;
            const float ot = top ? 0.0f : qo[c + sz], dt_ = top ? 0.0f : dq[c + sz];
            const float uw = Um[l.u0 + (off)k * (d.nx + 1)];
            const float vs = Vm[c], vn = Vm[c + sj], wt = Wm[c];
            const float rj = Rj[c], rdj = Rdj[c];
            const float qw = qo[c - 1], qs = qo[c - sj], qn = qo[c + sj];
            // the upwind fluxes of both half steps and flux3 with t_factor 1
            const float uxw = up2(uw, qw, o0, dq[c - 1], d0);
            const float hxw = face_h<H>(uw, qin, c, 1);
            const float uxe = __shfl_down(uxw, 1), hxe = __shfl_down(hxw, 1);
            const float ue = __shfl_down(uw, 1), qe = __shfl_down(o0, 1);
            const float uys = up2(vs, qs, o0, dq[c - sj], d0);
            const float uyn = up2(vn, o0, qn, d0, dq[c + sj]);
            const float hys = face_h<H>(vs, qin, c, sj);
            const float hyn = face_h<H>(vn, qin, c + sj, sj);
            float uzt;
            if (top) {  // :169, :184
                uzt = 0.5f * o0 * wt;
                uzt = uzt + 0.5f * d0 * wt;
            } else {
                uzt = up2(wt, o0, ot, d0, dt_);
            }
            const float hzt = top ? zw[2] * wt : face_vw<V>(wt, zw, k + 1, d.nz);
            if (l.cell) {
                // :75-78 the upwind-only value
                const float temp = flux_update(o0, uxw, uxe, uys, uyn, uzb, uzt, rj, rdj);
                // :80-85 merged fluxes
                const float mxw = hxw - uxw, mxe = hxe - uxe, mys = hys - uys, myn = hyn - uyn,
                            mzb = hzb - uzb, mzt = hzt - uzt;
                // :88-96
                float fin = -((fmn(0.0f, mxe) - fmx(0.0f, mxw)) + (fmn(0.0f, myn) - fmx(0.0f, mys))) / rj;
                fin = fin - (fmn(0.0f, mzt) - fmx(0.0f, mzb)) / rdj;
                float fout = ((fmx(0.0f, mxe) - fmn(0.0f, mxw)) + (fmx(0.0f, myn) - fmn(0.0f, mys))) / rj;
                fout = fout + (fmx(0.0f, mzt) - fmn(0.0f, mzb)) / rdj;
                // :99-121 bounds from the upwind neighbours (unscaled winds)
                float qmax = o0, qmin = o0;
                if (uw > 0.0f) {
                    qmax = fmx(qw, qmax);
                    qmin = fmn(qw, qmin);
                } else if (ue < 0.0f) {
                    qmax = fmx(qe, qmax);
                    qmin = fmn(qe, qmin);
                }
                if (vs > 0.0f) {
                    qmax = fmx(qs, qmax);
                    qmin = fmn(qs, qmin);
                } else if (vn < 0.0f) {
                    qmax = fmx(qn, qmax);
                    qmin = fmn(qn, qmin);
                }
                if (k > 0 && wb > 0.0f) {
                    qmax = fmx(ob, qmax);
                    qmin = fmn(ob, qmin);
                }
                if (!top && wt < 0.0f) {
                    qmax = fmx(ot, qmax);
                    qmin = fmn(ot, qmin);
                }
                // :124-125
                float si = 1.0f, so = 1.0f;
                if (fin > (qmax - temp)) si = fmx(0.0f, (qmax - temp) / (fin + 1.E-15f));
                if (fout > (temp - qmin)) so = fmx(0.0f, (temp - qmin) / (fout + 1.E-15f));
                put(b.sc[s], c, make_float2(si, so));
            }
            uzb = uzt;
            hzb = hzt;
            wb = wt;
            ob = o0;
            o0 = ot;
            d0 = dt_;
#pragma unroll
            for (int m = 0; m < 5; m++) zw[m] = zw[m + 1];
            zw[5] = k + 4 < d.nz ? qin[c + 4 * sz] : 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------
// the limited fluxes and the final update, on the tile
// ---------------------------------------------------------------------------
template <int H, int V>
__global__ void __launch_bounds__(64 * RB) k_std_final(StdGeom g, StdBatch b, Region r)
{
    const Dims d = g.d;
    const Lane l = lane_of(d, r);
    if (l.j > r.j_hi) return;
    const G<float> Um{g.Um}, Vm{g.Vm}, Wm{g.Wm}, Rj{g.Rj}, Rdj{g.Rdj};
    const off sj = (off)d.nx * d.nz, sz = d.nx;
    for (int s = l.s0; s < min(l.s1, b.n); s++) {
        const G<float> qin{b.qin[s]}, qo{b.qold[s]}, dq{b.dq[s]};
        const G<float2> sc{b.sc[s]};
        float zw[6], w7[7];
        load_window7(w7, qin, l.c0, l.k0, d.nz, sz);
#pragma unroll
        for (int m = 0; m < 6; m++) zw[m] = w7[m + 1];
        const off ck = l.c0 + (off)l.k0 * sz;
        float o0 = qo[ck], d0 = dq[ck];
        float2 s0 = sc[ck];  // this cell's scales
        float Fzb = 0.0f;    // through the face under the level
        if (l.k0 > 0) {
            const float wb = Wm[ck - sz];
            const float uzb = up2(wb, qo[ck - sz], o0, dq[ck - sz], d0);
            Fzb = limit_face(face_under<V>(wb, w7, l.k0, d.nz) - uzb, s0, sc[ck - sz]) + uzb;
        }
#pragma unroll 1
        for (int k = l.k0; k < l.k1; k++) {
            const off c = l.c0 + (off)k * sz;
            const bool top = k == d.nz - 1;
            const float ot = top ? 0.0f : qo[c + sz], dt_ = top ? 0.0f : dq[c + sz];
            const float2 st = top ? make_float2(0.0f, 0.0f) : sc[c + sz];
            const float uw = Um[l.u0 + (off)k * (d.nx + 1)];
            const float vs = Vm[c], vn = Vm[c + sj], wt = Wm[c];
            const float rj = Rj[c], rdj = Rdj[c];
            // x: the west face, normalised with this cell (positive side) and
            // the one to the west; the east face is the next lane's
            const float uxw = up2(uw, qo[c - 1], o0, dq[c - 1], d0);
            const float Fxw = limit_face(face_h<H>(uw, qin, c, 1) - uxw, s0, sc[c - 1]) + uxw;
            const float Fxe = __shfl_down(Fxw, 1);
            const float uys = up2(vs, qo[c - sj], o0, dq[c - sj], d0);
            const float uyn = up2(vn, o0, qo[c + sj], d0, dq[c + sj]);
            const float Fys = limit_face(face_h<H>(vs, qin, c, sj) - uys, s0, sc[c - sj]) + uys;
            const float Fyn = limit_face(face_h<H>(vn, qin, c + sj, sj) - uyn, sc[c + sj], s0) + uyn;
            float uzt, Fzt;
            if (top) {  // merged but not limited (:139-143 stop at kme)
                uzt = 0.5f * o0 * wt;
                uzt = uzt + 0.5f * d0 * wt;
                Fzt = (zw[2] * wt - uzt) + uzt;
            } else {
                uzt = up2(wt, o0, ot, d0, dt_);
                Fzt = limit_face(face_vw<V>(wt, zw, k + 1, d.nz) - uzt, st, s0) + uzt;
            }
            if (l.cell) put(b.out[s], c, flux_update(o0, Fxw, Fxe, Fys, Fyn, Fzb, Fzt, rj, rdj));
            Fzb = Fzt;
            o0 = ot;
            d0 = dt_;
            s0 = st;
#pragma unroll
            for (int m = 0; m < 5; m++) zw[m] = zw[m + 1];
            zw[5] = k + 4 < d.nz ? qin[c + 4 * sz] : 0.0f;
        }
    }
}

// qold into out outside the tile, and into dq outside the tile plus one ring
// The columns outside the tile, numbered: the rows below and above the
// tile (whole rows), then the columns west and east of it on the tile's rows.
__host__ __device__ inline long long ring_columns(const StdGeom &g)
{
    const long long nx = g.d.nx, ny = g.d.ny, tr = g.jte - g.jts + 1;
    return (long long)g.jts * nx + (ny - 1 - g.jte) * nx + tr * (g.its + (nx - 1 - g.ite));
}

__global__ void __launch_bounds__(256) k_std_ring(StdGeom g, StdBatch b, int with_dq)
{
    const Dims d = g.d;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (t >= ring_columns(g)) return;
    const long long below = (long long)g.jts * d.nx, above = (long long)(d.ny - 1 - g.jte) * d.nx;
    const int side = g.its + (d.nx - 1 - g.ite);
    int i, j;
    if (t < below) {
        j = (int)(t / d.nx);
        i = (int)(t % d.nx);
    } else if (t < below + above) {
        j = g.jte + 1 + (int)((t - below) / d.nx);
        i = (int)((t - below) % d.nx);
    } else {
        const long long u = t - below - above;
        j = g.jts + (int)(u / side);
        const int m = (int)(u % side);
        i = m < g.its ? m : g.ite + 1 + (m - g.its);
    }
    const bool ext = i >= g.its - 1 && i <= g.ite + 1 && j >= g.jts - 1 && j <= g.jte + 1;
    const off c = (off)i + (off)d.nx * ((off)k + (off)d.nz * j);
    for (int s = 0; s < b.n; s++) {
        const float v = b.qold[s][c];
        b.out[s][c] = v;
        if (with_dq && !ext) b.dq[s][c] = v;
    }
}

dim3 cells_grid(const Dims &d) { return dim3((d.nx + 63) / 64, (d.nz + 3) / 4, d.ny); }

// levels per chunk: at most 16, in equal chunks (two chunks of 15 for 30
// levels measured best on the 1024x1024x30 domain: 8-level chunks cost 3%,
// whole columns 7%, both from L2 reuse between the rows of tiles in flight)
int chunk_levels(int nz)
{
    const int n = std::max(1, (nz + 15) / 16);
    return (nz + n - 1) / n;
}

// species per block of the limiter kernels: one species per block keeps
// the blocks short, so that neighbouring tiles run in step and share their
// stencil rows in L2 (measured 5% faster than a block looping over ten)
constexpr int SPECIES_PER_BLOCK = 1;

Region with_blocks(Region r, int nz, int nspecies)
{
    r.nbx = (r.i_hi - r.i_lo + 1 + CW - 1) / CW;
    r.nby = (r.j_hi - r.j_lo + 1 + RB - 1) / RB;
    r.nbz = (nz + r.kch - 1) / r.kch;
    r.sg = std::min(nspecies, SPECIES_PER_BLOCK);
    return r;
}

dim3 march_grid(const Region &r, int nspecies)
{
    const unsigned groups = (unsigned)((nspecies + r.sg - 1) / r.sg);
    return dim3((unsigned)r.nbx * r.nby * r.nbz * groups);
}

Region tile_of(const StdGeom &g, int nspecies)
{
    return with_blocks({g.its, g.ite, g.jts, g.jte, chunk_levels(g.d.nz), 0, 0, 0, 0}, g.d.nz,
                       nspecies);
}
Region ext_of(const StdGeom &g, int nspecies)
{
    return with_blocks({g.its - 1, g.ite + 1, g.jts - 1, g.jte + 1, chunk_levels(g.d.nz), 0, 0, 0, 0},
                       g.d.nz, nspecies);
}

template <int H, int V>
void launch_std_orders(hipStream_t s, int what, const StdGeom &g, float t, int mode,
                       const StdBatch &b)
{
    const dim3 block(64, RB);
    if (what == 0) {
        const Region r = mode == 0 ? tile_of(g, b.n) : ext_of(g, b.n);
        const dim3 cg((r.i_hi - r.i_lo + 64) / 64, (g.d.nz + 3) / 4, r.j_hi - r.j_lo + 1);
        if (mode == 0) hipLaunchKernelGGL((k_std_stage_cell<H, V, 0>), cg, dim3(64, 4), 0, s, g, t, b, r);
        if (mode == 1) hipLaunchKernelGGL((k_std_stage_cell<H, V, 1>), cg, dim3(64, 4), 0, s, g, t, b, r);
        if (mode == 2) hipLaunchKernelGGL((k_std_stage_cell<H, V, 2>), cg, dim3(64, 4), 0, s, g, t, b, r);
    } else if (what == 1) {
        const Region r = ext_of(g, b.n);
        hipLaunchKernelGGL((k_std_scale<H, V>), march_grid(r, b.n), block, 0, s, g, b, r);
    } else {
        const Region r = tile_of(g, b.n);
        hipLaunchKernelGGL((k_std_final<H, V>), march_grid(r, b.n), block, 0, s, g, b, r);
    }
}

template <int H>
void launch_std_v(hipStream_t s, int what, const StdGeom &g, float t, int mode, const StdBatch &b)
{
    if (g.vorder == 1) launch_std_orders<H, 1>(s, what, g, t, mode, b);
    else if (g.vorder == 3) launch_std_orders<H, 3>(s, what, g, t, mode, b);
    else launch_std_orders<H, 5>(s, what, g, t, mode, b);
}

void launch_std(hipStream_t s, int what, const StdGeom &g, float t, int mode, const StdBatch &b)
{
    if (g.horder == 1) launch_std_v<1>(s, what, g, t, mode, b);
    else if (g.horder == 3) launch_std_v<3>(s, what, g, t, mode, b);
    else launch_std_v<5>(s, what, g, t, mode, b);
}

}  // namespace

void launch_std_wind(hipStream_t s, Dims d, float dx, float dt, int adv_den,
                     const DomainFields &f, float *Um, float *Vm, float *Wm, float *Rj,
                     float *Rdj)
{
    hipLaunchKernelGGL(k_std_wind, cells_grid(d), dim3(64, 4), 0, s, d, dx, dt, adv_den, f, Um, Vm,
                       Wm, Rj, Rdj);
}

void launch_std_stage(hipStream_t s, const StdGeom &g, float t_factor, int mode,
                      const StdBatch &b)
{
    launch_std(s, 0, g, t_factor, mode, b);
}

void launch_std_scale(hipStream_t s, const StdGeom &g, const StdBatch &b)
{
    launch_std(s, 1, g, 1.0f, 0, b);
}

void launch_std_final(hipStream_t s, const StdGeom &g, const StdBatch &b)
{
    launch_std(s, 2, g, 1.0f, 0, b);
}

void launch_std_ring(hipStream_t s, const StdGeom &g, const StdBatch &b, int with_dq)
{
    const long long cols = ring_columns(g);
    if (cols <= 0) return;
    hipLaunchKernelGGL(k_std_ring, dim3((unsigned)((cols + 255) / 256), g.d.nz), dim3(256), 0, s, g,
                       b, with_dq);
}

}  // namespace icaradv
