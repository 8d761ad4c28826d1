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
// The stage, scale and final kernels march up the columns of LDS-staged
// tiles, one species per block, and compute each face flux once (see Tile
// below).  Every evaluation is the same IEEE FP32
// expression in the Fortran's order (this unit is compiled with
// -ffp-contract=off), so the results are the bits of a strict build of the
// reference.  Fortran MAX/MIN of two reals are the selects a > b ? a : b and
// a < b ? a : b.
#include <algorithm>

#include "common.hpp"

namespace icaradv {

namespace {

// Element offsets are 32-bit and the byte offsets of float arrays are
// formed in 32 bits (the host limits a domain to 2^30 cells), so that a load
// is a uniform base in scalar registers plus one 32-bit vector offset.
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

// (scale_in, scale_out) pairs: one 64-bit load, its byte offset formed in
// 64 bits (8-byte elements of up to 2^30 cells)
template <>
struct G<float2> {
    const float2 *p;
    __device__ __forceinline__ float2 operator[](off i) const
    {
        typedef __attribute__((address_space(1))) const uint64_t gcu64;
        const uint64_t v = ((gcu64 *)p)[i];
        return make_float2(__uint_as_float((uint32_t)v), __uint_as_float((uint32_t)(v >> 32)));
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
    typedef __attribute__((address_space(1))) uint64_t gu64;
    ((gu64 *)p)[i] = (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32);
}

// A block of cells [i_lo, i_hi] x [j_lo, j_hi] (0-based), all levels, cut
// into tiles and chunks of kch levels.
struct Region {
    int i_lo, i_hi, j_lo, j_hi, kch;
    int nbx, nby, nbz;  // tiles along i and j, chunks along k
    int ns;             // species of the launch, one per block
};

__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }

// x / C for the constants C = 12 and 60 of flux3, with R = 1/C rounded: the
// quotient estimate q = x*R, its remainder x - q*C (exact in one FMA) and
// q + remainder*R rounded once.  Compared with the IEEE division over every
// float for both constants, this returns the IEEE quotient for each x with
// |x| >= 2^-124 and for +-0 (whose sign copysign restores); smaller x (whose
// quotient is near or below the subnormal range), infinities and NaNs take
// the IEEE division, a branch the lanes of a wave skip unless one needs it.
template <int C>
__device__ __forceinline__ float div_const(float x)
{
    constexpr float R = 1.0f / (float)C;
    const uint32_t a = __float_as_uint(x) & 0x7fffffffu;
    if (__builtin_expect(a - 1u < 0x017fffffu || a >= 0x7f800000u, 0)) return x / (float)C;
    const float q = x * R;
    const float e = __builtin_fmaf(-q, (float)C, x);
    return __builtin_copysignf(__builtin_fmaf(e, R, q), x);
}

// flux3's face formulas (advect.f90:150-178): m3..p2 are q at offsets -3..+2
// from the cell on the positive side of the face.
__device__ __forceinline__ float face_up(float u, float m1, float p0)
{
    return ((u + fabsf(u)) * m1 + (u - fabsf(u)) * p0) / 2.0f;
}

__device__ __forceinline__ float face_o3(float u, float m2, float m1, float p0, float p1)
{
    float f = 7.0f * (p0 + m1) - (p1 + m2);
    f = div_const<12>(u * f);
    f = f - div_const<12>(fabsf(u)) * (3.0f * (p0 - m1) - (p1 - m2));
    return f;
}

__device__ __forceinline__ float face_o5(float u, float m3, float m2, float m1, float p0,
                                         float p1, float p2)
{
    float f = 37.0f * (p0 + m1) - 8.0f * (p1 + m2) + (p2 + m3);
    f = div_const<60>(u * f);
    f = f - div_const<60>(fabsf(u)) * (10.0f * (p0 - m1) - 5.0f * (p1 - m2) + (p2 - m3));
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

constexpr int CW = 63;  // columns a wave updates (lane 63 supplies a face only)

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
// Column marches over tiles staged in LDS
// ---------------------------------------------------------------------------
// A block covers CW columns by TY rows of the region, one wave per row, plus
// a wave for row TY that supplies the y faces north of the tile's last row.
// Its lanes march up a chunk of levels together.  At each level the block
// stages the planes its stencils read in LDS: each lane its own cell, and a
// few waves the halo around the tile.  Then every face is computed once:
//   x  on the west side of each lane's column (the east face is the next
//      lane's; lane 63 only supplies it),
//   y  on the south side of each row (the north face is the next row's,
//      passed through LDS after a second barrier),
//   z  above each level (the face under a level is carried in registers).
// Blocks take one species each, the species fastest in the block order, so
// that the species of a tile run together on one XCD and read the tile's
// winds and divisors from its L2.
template <int H, int TY>
struct Tile {
    static constexpr int HH = (H + 1) / 2;         // stencil half width of flux3
    static constexpr int QR = TY + 2 * HH;         // qin plane: rows j0-HH .. j0+TY+HH-1
    static constexpr int QC = CW + 2 * HH;         //   columns i0-HH .. i0+62+HH
    static constexpr int PR = TY + 2;              // qold/dumb_q/scales: rows j0-1 .. j0+TY
    static constexpr int PC = 65;                  //   columns i0-1 .. i0+63
    static constexpr int NXH = (2 * HH - 1) * TY;  // qin cells west and east of the tile rows
    static constexpr int NT = 64 * (TY + 1);       // threads
    // halo roles: waves 0 .. 2HH-2 stage a qin row below or above the tile,
    // wave TY-2 the column west of the tile and wave TY-1 the row south of
    // it in the other planes, wave TY the qin cells west and east of the rows
    static_assert(TY - 2 > 2 * HH - 2 && NXH <= 128, "tile rows");
};

struct TLane {
    int L, w;         // lane; wave = row of the tile (w == TY: y faces only)
    int i, j;         // this lane's column and row
    int k0, k1, s;    // levels [k0, k1) and the species
    bool cell;        // the lane updates (i, j)
    off cb, ub;       // (i, j) at level 0 in cells and in the x-face array
    int pq, pp;       // (i, j) in the qin plane and in the other planes
    off hq[2];        // halo: up to two qin cells (level 0 offsets)
    int hqd[2];       //   their place in the qin plane, -1 if none
    bool hqv[2];      //   inside the domain
    off ho;           // halo: one cell of the other planes
    int hod;          //   its place, -1 if none
    bool hov;
};

template <int H, int TY>
__device__ __forceinline__ TLane tile_lane(const Dims &d, const Region &r)
{
    using T = Tile<H, TY>;
    constexpr int HH = T::HH;
    TLane l;
    unsigned t = xcd_block();
    l.s = (int)(t % (unsigned)r.ns);
    t /= (unsigned)r.ns;
    const int bz = (int)(t % (unsigned)r.nbz);
    t /= (unsigned)r.nbz;
    const int bx = (int)(t % (unsigned)r.nbx), by = (int)(t / (unsigned)r.nbx);
    l.L = (int)threadIdx.x & 63;
    l.w = (int)threadIdx.x >> 6;
    const int i0 = r.i_lo + bx * CW, j0 = r.j_lo + by * TY;
    l.i = i0 + l.L;
    l.j = j0 + l.w;
    l.k0 = bz * r.kch;
    l.k1 = min(d.nz, l.k0 + r.kch);
    l.cell = l.L < CW && l.w < TY && l.i <= r.i_hi && l.j <= r.j_hi;
    const bool dom = l.i < d.nx && l.j < d.ny;
    const off sj = (off)d.nx * (off)d.nz;
    // a lane outside the domain reads cell 0: its values reach only faces
    // that no updated cell uses
    l.cb = dom ? (off)l.i + sj * (off)l.j : 0u;
    l.ub = dom ? (off)l.i + (off)(d.nx + 1) * (off)d.nz * (off)l.j : 0u;
    l.pq = (HH + l.w) * T::QC + HH + l.L;
    l.pp = (1 + l.w) * T::PC + 1 + l.L;
    l.hqd[0] = l.hqd[1] = -1;
    l.hq[0] = l.hq[1] = 0u;
    l.hqv[0] = l.hqv[1] = false;
    l.hod = -1;
    l.ho = 0u;
    l.hov = false;
    auto at = [&](int ii, int jj, bool &v) -> off {
        v = ii >= 0 && ii < d.nx && jj >= 0 && jj < d.ny;
        return v ? (off)ii + sj * (off)jj : 0u;
    };
    if (l.w < 2 * HH - 1) {  // a qin row below (e < HH) or above the tile
        const int e = l.w;
        const int pr = e < HH ? e : TY + HH + 1 + (e - HH);
        l.hq[0] = at(l.i, j0 - HH + pr, l.hqv[0]);
        l.hqd[0] = pr * T::QC + HH + l.L;
    } else if (l.w == TY - 2) {  // column i0-1 of the rows j0 .. j0+TY-1
        if (l.L < TY) {
            l.ho = at(i0 - 1, j0 + l.L, l.hov);
            l.hod = (1 + l.L) * T::PC;
        }
    } else if (l.w == TY - 1) {  // row j0-1 of the columns i0 .. i0+63
        l.ho = at(l.i, j0 - 1, l.hov);
        l.hod = 1 + l.L;
    } else if (l.w == TY) {  // qin columns i0-HH .. i0-1 and i0+64 .. i0+62+HH
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int e = l.L + 64 * u;
            if (e < T::NXH) {
                const int rr = e / (2 * HH - 1), cc = e % (2 * HH - 1);
                const int q = cc < HH ? cc : 64 + cc;
                l.hq[u] = at(i0 - HH + q, j0 + rr, l.hqv[u]);
                l.hqd[u] = (HH + rr) * T::QC + q;
            }
        }
    }
    return l;
}

// the halo values of level k: qin (Q) and the first NP of qold, dumb_q and
// the scales
struct THalo {
    float q[2], o, dq;
    float2 sc;
};

template <bool Q, int NP>
__device__ __forceinline__ THalo tile_halo(const TLane &l, G<float> qin, G<float> qo, G<float> dq,
                                           G<float2> sc, int k, off sz)
{
    THalo h;
    const off kk = (off)k * sz;
#pragma unroll
    for (int u = 0; u < 2; u++) h.q[u] = Q && l.hqv[u] ? qin[l.hq[u] + kk] : 0.0f;
    h.o = h.dq = 0.0f;
    h.sc = make_float2(0.0f, 0.0f);
    if (NP >= 1 && l.hov) {
        h.o = qo[l.ho + kk];
        if (NP >= 2) h.dq = dq[l.ho + kk];
        if (NP >= 3) h.sc = sc[l.ho + kk];
    }
    return h;
}

template <bool Q, int NP>
__device__ __forceinline__ void stage_halo(const TLane &l, const THalo &h, float *sq, float *so,
                                           float *sd, float *ssi, float *sso)
{
    if (Q) {
#pragma unroll
        for (int u = 0; u < 2; u++)
            if (l.hqd[u] >= 0) sq[l.hqd[u]] = h.q[u];
    }
    if (NP >= 1 && l.hod >= 0) {
        so[l.hod] = h.o;
        if (NP >= 2) sd[l.hod] = h.dq;
        if (NP >= 3) {
            ssi[l.hod] = h.sc.x;
            sso[l.hod] = h.sc.y;
        }
    }
}

// the winds and divisors of a cell at one level (the last wave needs only
// the south face's wind)
struct TWind {
    float uw, vs, wt, rj, rdj;
};

__device__ __forceinline__ TWind tile_wind(const TLane &l, const StdGeom &g, int k, bool row)
{
    const G<float> Um{g.Um}, Vm{g.Vm}, Wm{g.Wm}, Rj{g.Rj}, Rdj{g.Rdj};
    const off c = l.cb + (off)k * (off)g.d.nx;
    TWind f;
    f.vs = Vm[c];
    f.uw = f.wt = 0.0f;
    f.rj = f.rdj = 1.0f;
    if (row) {
        f.uw = Um[l.ub + (off)k * (off)(g.d.nx + 1)];
        f.wt = Wm[c];
        f.rj = Rj[c];
        f.rdj = Rdj[c];
    }
    return f;
}

// Each iteration stages level k from registers, then issues the loads of
// what the next iteration stages and computes with (the halo of level k+1,
// the winds of level k+1, the lane's own values of level k+2), so that no
// load is waited for in the iteration that issues it.

// One RK3 stage / Euler step over tiles: flux3 of the stage state with the
// winds times t applied to the base state (MODE 0 on the tile; MODE 1 the
// same plus the dumb_q half step on the tile and one ring; MODE 2 dumb_q
// alone on that region)
template <int H, int V, int MODE, int TY>
__global__ void __launch_bounds__(64 * (TY + 1))
    k_std_stage(StdGeom g, float t, StdBatch b, Region r)
{
    using T = Tile<H, TY>;
    constexpr bool FL = MODE <= 1, DQ = MODE >= 1;
    __shared__ float sq[FL ? T::QR * T::QC : 1];
    __shared__ float so[DQ ? T::PR * T::PC : 1];
    __shared__ float sy[(TY + 1) * 64], sh[DQ ? (TY + 1) * 64 : 1];
    const Dims d = g.d;
    const TLane l = tile_lane<H, TY>(d, r);
    const bool row = l.w < TY;
    const bool in_tile = l.i >= g.its && l.i <= g.ite && l.j >= g.jts && l.j <= g.jte;
    const bool put_fl = FL && l.cell && (MODE == 0 || in_tile);
    const G<float> Wm{g.Wm};
    const G<float> qin{b.qin[l.s]}, qo{b.qold[l.s]};
    const G<float2> nosc{nullptr};
    const off sz = d.nx;
    float zw[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, w7[7];
    if (FL) {
        load_window7(w7, qin, l.cb, l.k0, d.nz, sz);
#pragma unroll
        for (int m = 0; m < 6; m++) zw[m] = w7[m + 1];
    }
    const off ck = l.cb + (off)l.k0 * sz;
    float o0 = qo[ck], ot = l.k0 + 1 < d.nz ? qo[ck + sz] : 0.0f;  // qold here and above
    float fzb = 0.0f, hzb = 0.0f;  // the faces under the level
    if (row && l.k0 > 0) {
        const float wb = Wm[ck - sz];
        if (FL) fzb = face_under<V>(wb * t, w7, l.k0, d.nz);
        if (DQ) hzb = half_up(wb, qo[ck - sz], o0);
    }
    THalo h = tile_halo<FL, DQ ? 1 : 0>(l, qin, qo, qo, nosc, l.k0, sz);
    TWind f = tile_wind(l, g, l.k0, row);
#pragma unroll 1
    for (int k = l.k0; k < l.k1; k++) {
        const off c = l.cb + (off)k * sz;
        const bool top = k == d.nz - 1;
        if (FL) sq[l.pq] = zw[2];
        if (DQ) so[l.pp] = o0;
        stage_halo<FL, DQ ? 1 : 0>(l, h, sq, so, nullptr, nullptr, nullptr);
        const float on = k + 2 < d.nz ? qo[c + 2 * sz] : 0.0f;
        TWind fn = f;
        if (k + 1 < l.k1) {
            h = tile_halo<FL, DQ ? 1 : 0>(l, qin, qo, qo, nosc, k + 1, sz);
            fn = tile_wind(l, g, k + 1, row);
        }
        __syncthreads();
        // flux3 with the winds times t (advect.f90:285-287) and the first
        // upwind half step (adv_fluxcorr.f90:158-175) through the west face,
        // the face above the level and the south face
        float fxw = 0.0f, fzt = 0.0f, hxw = 0.0f, hzt = 0.0f, fys = 0.0f, hys = 0.0f;
        if (row) {
            if (FL) {
                fxw = face_h<H>(f.uw * t, sq, l.pq, 1);
                fzt = top ? zw[2] * (f.wt * t) : face_vw<V>(f.wt * t, zw, k + 1, d.nz);  // :253
            }
            if (DQ) {
                hxw = half_up(f.uw, so[l.pp - 1], o0);
                hzt = top ? 0.5f * o0 * f.wt : half_up(f.wt, o0, ot);
            }
        }
        const float fxe = FL ? __shfl_down(fxw, 1) : 0.0f;
        const float hxe = DQ ? __shfl_down(hxw, 1) : 0.0f;
        if (FL) {
            fys = face_h<H>(f.vs * t, sq, l.pq, T::QC);
            sy[l.w * 64 + l.L] = fys;
        }
        if (DQ) {
            hys = half_up(f.vs, so[l.pp - T::PC], o0);
            sh[l.w * 64 + l.L] = hys;
        }
        __syncthreads();
        const int n = (l.w + 1) * 64 + l.L;
        if (put_fl) put(b.out[l.s], c, flux_update(o0, fxw, fxe, fys, sy[n], fzb, fzt, f.rj, f.rdj));
        if (DQ && l.cell)
            put(b.dq[l.s], c, flux_update(o0, hxw, hxe, hys, sh[n], hzb, hzt, f.rj, f.rdj));
        fzb = fzt;
        hzb = hzt;
        o0 = ot;
        ot = on;
        f = fn;
        if (FL) {
#pragma unroll
            for (int m = 0; m < 5; m++) zw[m] = zw[m + 1];
            zw[5] = k + 4 < d.nz ? qin[c + 4 * sz] : 0.0f;
        }
    }
}

// WRF_flux_corr up to scale_in / scale_out (the region is the tile plus one
// ring of cells)
template <int H, int V, int TY>
__global__ void __launch_bounds__(64 * (TY + 1)) k_std_scale(StdGeom g, StdBatch b, Region r)
{
    using T = Tile<H, TY>;
    __shared__ float sq[T::QR * T::QC];
    __shared__ float so[T::PR * T::PC], sd[T::PR * T::PC];
    __shared__ float syu[(TY + 1) * 64], syh[(TY + 1) * 64], syv[(TY + 1) * 64];
    const Dims d = g.d;
    const TLane l = tile_lane<H, TY>(d, r);
    const bool row = l.w < TY;
    const G<float> Wm{g.Wm};
    const G<float> qin{b.qin[l.s]}, qo{b.qold[l.s]}, dq{b.dq[l.s]};
    const G<float2> nosc{nullptr};
    const off sz = d.nx;
    float zw[6], w7[7];
    load_window7(w7, qin, l.cb, l.k0, d.nz, sz);
#pragma unroll
    for (int m = 0; m < 6; m++) zw[m] = w7[m + 1];
    const off ck = l.cb + (off)l.k0 * sz;
    // qold below, here and above; dumb_q here and above
    float ob = 0.0f, o0 = qo[ck], d0 = dq[ck], ot = 0.0f, dt_ = 0.0f;
    if (l.k0 + 1 < d.nz) {
        ot = qo[ck + sz];
        dt_ = dq[ck + sz];
    }
    float uzb = 0.0f, hzb = 0.0f, wb = 0.0f;  // the face under the level
    if (row && l.k0 > 0) {
        wb = Wm[ck - sz];
        ob = qo[ck - sz];
        uzb = up2(wb, ob, o0, dq[ck - sz], d0);
        hzb = face_under<V>(wb, w7, l.k0, d.nz);
    }
    THalo h = tile_halo<true, 2>(l, qin, qo, dq, nosc, l.k0, sz);
    TWind f = tile_wind(l, g, l.k0, row);
#pragma unroll 1
    for (int k = l.k0; k < l.k1; k++) {
        const off c = l.cb + (off)k * sz;
        const bool top = k == d.nz - 1;
        sq[l.pq] = zw[2];
        so[l.pp] = o0;
        sd[l.pp] = d0;
        stage_halo<true, 2>(l, h, sq, so, sd, nullptr, nullptr);
        float on = 0.0f, dn = 0.0f;
        if (k + 2 < d.nz) {
            on = qo[c + 2 * sz];
            dn = dq[c + 2 * sz];
        }
        TWind fn = f;
        if (k + 1 < l.k1) {
            h = tile_halo<true, 2>(l, qin, qo, dq, nosc, k + 1, sz);
            fn = tile_wind(l, g, k + 1, row);
        }
        __syncthreads();
        // the upwind fluxes of both half steps and flux3 with t_factor 1
        float uxw = 0.0f, hxw = 0.0f, qw = 0.0f, qn = 0.0f, uzt = 0.0f, hzt = 0.0f;
        if (row) {
            qw = so[l.pp - 1];
            qn = so[l.pp + T::PC];
            uxw = up2(f.uw, qw, o0, sd[l.pp - 1], d0);
            hxw = face_h<H>(f.uw, sq, l.pq, 1);
            if (top) {  // :169, :184
                uzt = 0.5f * o0 * f.wt;
                uzt = uzt + 0.5f * d0 * f.wt;
            } else {
                uzt = up2(f.wt, o0, ot, d0, dt_);
            }
            hzt = top ? zw[2] * f.wt : face_vw<V>(f.wt, zw, k + 1, d.nz);
        }
        const float qs = so[l.pp - T::PC];
        const float uxe = __shfl_down(uxw, 1), hxe = __shfl_down(hxw, 1);
        const float ue = __shfl_down(f.uw, 1), qe = __shfl_down(o0, 1);
        const float uys = up2(f.vs, qs, o0, sd[l.pp - T::PC], d0);
        const float hys = face_h<H>(f.vs, sq, l.pq, T::QC);
        syu[l.w * 64 + l.L] = uys;
        syh[l.w * 64 + l.L] = hys;
        syv[l.w * 64 + l.L] = f.vs;
        __syncthreads();
        if (l.cell) {
            const int n = (l.w + 1) * 64 + l.L;
            const float uyn = syu[n], hyn = syh[n], vn = syv[n];
            // :75-78 the upwind-only value
            const float temp = flux_update(o0, uxw, uxe, uys, uyn, uzb, uzt, f.rj, f.rdj);
            // :80-85 merged fluxes
            const float mxw = hxw - uxw, mxe = hxe - uxe, mys = hys - uys, myn = hyn - uyn,
                        mzb = hzb - uzb, mzt = hzt - uzt;
            // :88-96
            float fin = -((fmn(0.0f, mxe) - fmx(0.0f, mxw)) + (fmn(0.0f, myn) - fmx(0.0f, mys))) / f.rj;
            fin = fin - (fmn(0.0f, mzt) - fmx(0.0f, mzb)) / f.rdj;
            float fout = ((fmx(0.0f, mxe) - fmn(0.0f, mxw)) + (fmx(0.0f, myn) - fmn(0.0f, mys))) / f.rj;
            fout = fout + (fmx(0.0f, mzt) - fmn(0.0f, mzb)) / f.rdj;
            // :99-121 bounds from the upwind neighbours (unscaled winds)
            float qmax = o0, qmin = o0;
            if (f.uw > 0.0f) {
                qmax = fmx(qw, qmax);
                qmin = fmn(qw, qmin);
            } else if (ue < 0.0f) {
                qmax = fmx(qe, qmax);
                qmin = fmn(qe, qmin);
            }
            if (f.vs > 0.0f) {
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
            if (!top && f.wt < 0.0f) {
                qmax = fmx(ot, qmax);
                qmin = fmn(ot, qmin);
            }
            // :124-125
            float si = 1.0f, so_ = 1.0f;
            if (fin > (qmax - temp)) si = fmx(0.0f, (qmax - temp) / (fin + 1.E-15f));
            if (fout > (temp - qmin)) so_ = fmx(0.0f, (temp - qmin) / (fout + 1.E-15f));
            put(b.sc[l.s], c, make_float2(si, so_));
        }
        uzb = uzt;
        hzb = hzt;
        wb = f.wt;
        ob = o0;
        o0 = ot;
        d0 = dt_;
        ot = on;
        dt_ = dn;
        f = fn;
#pragma unroll
        for (int m = 0; m < 5; m++) zw[m] = zw[m + 1];
        zw[5] = k + 4 < d.nz ? qin[c + 4 * sz] : 0.0f;
    }
}

// the limited fluxes and the final update, on the tile
template <int H, int V, int TY>
__global__ void __launch_bounds__(64 * (TY + 1)) k_std_final(StdGeom g, StdBatch b, Region r)
{
    using T = Tile<H, TY>;
    __shared__ float sq[T::QR * T::QC];
    __shared__ float so[T::PR * T::PC], sd[T::PR * T::PC];
    __shared__ float ssi[T::PR * T::PC], sso[T::PR * T::PC];  // scale_in, scale_out
    __shared__ float sy[(TY + 1) * 64];
    const Dims d = g.d;
    const TLane l = tile_lane<H, TY>(d, r);
    const bool row = l.w < TY;
    const G<float> Wm{g.Wm};
    const G<float> qin{b.qin[l.s]}, qo{b.qold[l.s]}, dq{b.dq[l.s]};
    const G<float2> sc{b.sc[l.s]};
    const off sz = d.nx;
    float zw[6], w7[7];
    load_window7(w7, qin, l.cb, l.k0, d.nz, sz);
#pragma unroll
    for (int m = 0; m < 6; m++) zw[m] = w7[m + 1];
    const off ck = l.cb + (off)l.k0 * sz;
    // qold, dumb_q and the scales of this cell and the one above
    float o0 = qo[ck], d0 = dq[ck], ot = 0.0f, dt_ = 0.0f;
    float2 s0 = sc[ck], st = make_float2(0.0f, 0.0f);
    if (l.k0 + 1 < d.nz) {
        ot = qo[ck + sz];
        dt_ = dq[ck + sz];
        st = sc[ck + sz];
    }
    float Fzb = 0.0f;  // through the face under the level
    if (row && l.k0 > 0) {
        const float wb = Wm[ck - sz];
        const float uzb = up2(wb, qo[ck - sz], o0, dq[ck - sz], d0);
        Fzb = limit_face(face_under<V>(wb, w7, l.k0, d.nz) - uzb, s0, sc[ck - sz]) + uzb;
    }
    THalo h = tile_halo<true, 3>(l, qin, qo, dq, sc, l.k0, sz);
    TWind f = tile_wind(l, g, l.k0, row);
#pragma unroll 1
    for (int k = l.k0; k < l.k1; k++) {
        const off c = l.cb + (off)k * sz;
        const bool top = k == d.nz - 1;
        sq[l.pq] = zw[2];
        so[l.pp] = o0;
        sd[l.pp] = d0;
        ssi[l.pp] = s0.x;
        sso[l.pp] = s0.y;
        stage_halo<true, 3>(l, h, sq, so, sd, ssi, sso);
        float on = 0.0f, dn = 0.0f;
        float2 sn = make_float2(0.0f, 0.0f);
        if (k + 2 < d.nz) {
            on = qo[c + 2 * sz];
            dn = dq[c + 2 * sz];
            sn = sc[c + 2 * sz];
        }
        TWind fn = f;
        if (k + 1 < l.k1) {
            h = tile_halo<true, 3>(l, qin, qo, dq, sc, k + 1, sz);
            fn = tile_wind(l, g, k + 1, row);
        }
        __syncthreads();
        // x: the west face, normalised with this cell (positive side) and
        // the one to the west; z: the face above the level
        float Fxw = 0.0f, Fzt = 0.0f;
        if (row) {
            const float uxw = up2(f.uw, so[l.pp - 1], o0, sd[l.pp - 1], d0);
            Fxw = limit_face(face_h<H>(f.uw, sq, l.pq, 1) - uxw, s0,
                             make_float2(ssi[l.pp - 1], sso[l.pp - 1])) + uxw;
            if (top) {  // merged but not limited (:139-143 stop at kme)
                float uzt = 0.5f * o0 * f.wt;
                uzt = uzt + 0.5f * d0 * f.wt;
                Fzt = (zw[2] * f.wt - uzt) + uzt;
            } else {
                const float uzt = up2(f.wt, o0, ot, d0, dt_);
                Fzt = limit_face(face_vw<V>(f.wt, zw, k + 1, d.nz) - uzt, st, s0) + uzt;
            }
        }
        const float Fxe = __shfl_down(Fxw, 1);
        // y: the south face, the north one is the next row's
        const int ps = l.pp - T::PC;
        const float uys = up2(f.vs, so[ps], o0, sd[ps], d0);
        const float Fys =
            limit_face(face_h<H>(f.vs, sq, l.pq, T::QC) - uys, s0, make_float2(ssi[ps], sso[ps])) + uys;
        sy[l.w * 64 + l.L] = Fys;
        __syncthreads();
        if (l.cell)
            put(b.out[l.s], c,
                flux_update(o0, Fxw, Fxe, Fys, sy[(l.w + 1) * 64 + l.L], Fzb, Fzt, f.rj, f.rdj));
        Fzb = Fzt;
        o0 = ot;
        d0 = dt_;
        s0 = st;
        ot = on;
        dt_ = dn;
        st = sn;
        f = fn;
#pragma unroll
        for (int m = 0; m < 5; m++) zw[m] = zw[m + 1];
        zw[5] = k + 4 < d.nz ? qin[c + 4 * sz] : 0.0f;
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

// Tile rows: 7 rows and the y-face wave make 8 waves per block, which pack
// a CU exactly at the kernels' occupancy (6 waves per SIMD for the limiter
// kernels: 3 blocks, 8 for the stages: 4 blocks).  9- and 16-wave blocks
// measured 20-35% slower for leaving wave slots empty.
constexpr int TILE_ROWS = 7;

// The region in tiles, one species per block.  Whole columns, unless the
// grid would have too few blocks to fill the GPU; then the columns are cut
// into chunks (each chunk recomputes the face under its first level).
Region tiles(int i_lo, int i_hi, int j_lo, int j_hi, int nz, int nspecies)
{
    Region r{};
    r.i_lo = i_lo;
    r.i_hi = i_hi;
    r.j_lo = j_lo;
    r.j_hi = j_hi;
    r.nbx = (i_hi - i_lo + 1 + CW - 1) / CW;
    r.nby = (j_hi - j_lo + 1 + TILE_ROWS - 1) / TILE_ROWS;
    r.ns = nspecies;
    r.kch = nz;
    while (r.kch > 4 && (long long)r.nbx * r.nby * r.ns * ((nz + r.kch - 1) / r.kch) < 1024)
        r.kch = (r.kch + 1) / 2;
    r.nbz = (nz + r.kch - 1) / r.kch;
    return r;
}

Region tile_of(const StdGeom &g, int nspecies)
{
    return tiles(g.its, g.ite, g.jts, g.jte, g.d.nz, nspecies);
}

Region ext_of(const StdGeom &g, int nspecies)
{
    return tiles(g.its - 1, g.ite + 1, g.jts - 1, g.jte + 1, g.d.nz, nspecies);
}

dim3 tile_grid(const Region &r) { return dim3((unsigned)r.nbx * r.nby * r.nbz * r.ns); }

template <int H, int V>
void launch_std_orders(hipStream_t s, int what, const StdGeom &g, float t, int mode,
                       const StdBatch &b)
{
    const dim3 tb(Tile<H, TILE_ROWS>::NT);
    if (what == 0) {
        const Region r = mode == 0 ? tile_of(g, b.n) : ext_of(g, b.n);
        if (mode == 0) hipLaunchKernelGGL((k_std_stage<H, V, 0, TILE_ROWS>), tile_grid(r), tb, 0, s, g, t, b, r);
        if (mode == 1) hipLaunchKernelGGL((k_std_stage<H, V, 1, TILE_ROWS>), tile_grid(r), tb, 0, s, g, t, b, r);
        if (mode == 2) hipLaunchKernelGGL((k_std_stage<H, V, 2, TILE_ROWS>), tile_grid(r), tb, 0, s, g, t, b, r);
    } else if (what == 1) {
        const Region r = ext_of(g, b.n);
        hipLaunchKernelGGL((k_std_scale<H, V, TILE_ROWS>), tile_grid(r), tb, 0, s, g, b, r);
    } else {
        const Region r = tile_of(g, b.n);
        hipLaunchKernelGGL((k_std_final<H, V, TILE_ROWS>), tile_grid(r), tb, 0, s, g, b, r);
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
