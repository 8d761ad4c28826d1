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
// Every kernel takes one thread per cell (64 along i by 4 along k per block,
// one row j per block column) and loops over a batch of species with the
// winds and divisors held in registers.  A face flux is recomputed by each
// cell that uses it; every evaluation is the same IEEE FP32 expression in the
// Fortran's order (this unit is compiled with -ffp-contract=off), so the
// results are the bits of a strict build of the reference.  Fortran MAX/MIN
// of two reals are the selects a > b ? a : b and a < b ? a : b.
#include "common.hpp"

namespace icaradv {

namespace {

typedef ptrdiff_t off;

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

// horizontal face whose positive-side cell is c, stride s
template <int H>
__device__ __forceinline__ float face_h(float u, const float *__restrict__ q, off c, off s)
{
    if constexpr (H == 1) {
        return face_up(u, q[c - s], q[c]);
    } else if constexpr (H == 3) {
        return face_o3(u, q[c - 2 * s], q[c - s], q[c], q[c + s]);
    } else {
        return face_o5(u, q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s]);
    }
}

// vertical face kf (1..nz-1) whose upper cell is c (advect.f90:180-241): the
// faces next to the ground and the top stay upwind, 5th order falls back to
// 3rd one face further in
template <int V>
__device__ __forceinline__ float face_v(float w, const float *__restrict__ q, off c, int kf,
                                        int nz, off s)
{
    if (V == 1 || kf == 1 || kf == nz - 1) return face_up(w, q[c - s], q[c]);
    if (V == 5 && kf >= 3 && kf <= nz - 3)
        return face_o5(w, q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s]);
    return face_o3(w, q[c - 2 * s], q[c - s], q[c], q[c + s]);
}

// one upwind half step of upwind_flux3 (adv_fluxcorr.f90:160-165)
__device__ __forceinline__ float half_up(float u, float m1, float p0)
{
    return 0.5f * ((u + fabsf(u)) * m1 + (u - fabsf(u)) * p0) / 2.0f;
}

// The winds of one cell's six faces and its two divisors.
struct CellWinds {
    float uw, ue, vs, vn, wb, wt, rj, rdj;
};

// One value per face of a cell: west, east, south, north, bottom, top.
struct Six {
    float xw, xe, ys, yn, zb, zt;
};

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

__device__ __forceinline__ CellWinds load_winds(const StdGeom &g, const Cell &p)
{
    const Dims &d = g.d;
    const off fw = (off)p.i + (off)(d.nx + 1) * ((off)p.k + (off)d.nz * p.j);
    const off sj = (off)d.nx * d.nz;
    CellWinds w;
    w.uw = g.Um[fw];
    w.ue = g.Um[fw + 1];
    w.vs = g.Vm[p.c];
    w.vn = g.Vm[p.c + sj];
    w.wb = p.k > 0 ? g.Wm[p.c - d.nx] : 0.0f;
    w.wt = g.Wm[p.c];
    w.rj = g.Rj[p.c];
    w.rdj = g.Rdj[p.c];
    return w;
}

// flux3 at the six faces of cell p from state q with the winds times t
// (advect.f90:285-287 scales the winds before flux3)
template <int H, int V>
__device__ __forceinline__ Six flux3_six(const CellWinds &w, float t, const float *__restrict__ q,
                                         const Cell &p, const Dims &d)
{
    const off c = p.c, sx = 1, sj = (off)d.nx * d.nz, sz = d.nx;
    Six f;
    f.xw = face_h<H>(w.uw * t, q, c, sx);
    f.xe = face_h<H>(w.ue * t, q, c + sx, sx);
    f.ys = face_h<H>(w.vs * t, q, c, sj);
    f.yn = face_h<H>(w.vn * t, q, c + sj, sj);
    f.zb = p.k == 0 ? 0.0f : face_v<V>(w.wb * t, q, c, p.k, d.nz, sz);               // :252
    f.zt = p.k == d.nz - 1 ? q[c] * (w.wt * t) : face_v<V>(w.wt * t, q, c + sz, p.k + 1, d.nz, sz);  // :253
    return f;
}

// The flux-form update of adv_std_advect3d (advect.f90:301-309).
__device__ __forceinline__ float flux_update(float q0, const Six &f, const CellWinds &w)
{
    float r = q0 - ((f.xe - f.xw) + (f.yn - f.ys)) / w.rj;
    r = r - (f.zt - f.zb) / w.rdj;
    return r;
}

// The first upwind half step of upwind_flux3 at the six faces (from the base
// state qo, unscaled winds).
__device__ __forceinline__ Six upwind_half_six(const CellWinds &w, const float *__restrict__ qo,
                                               const Cell &p, const Dims &d)
{
    const off c = p.c, sj = (off)d.nx * d.nz, sz = d.nx;
    Six f;
    f.xw = half_up(w.uw, qo[c - 1], qo[c]);
    f.xe = half_up(w.ue, qo[c], qo[c + 1]);
    f.ys = half_up(w.vs, qo[c - sj], qo[c]);
    f.yn = half_up(w.vn, qo[c], qo[c + sj]);
    f.zb = p.k == 0 ? 0.0f : half_up(w.wb, qo[c - sz], qo[c]);              // :168
    f.zt = p.k == d.nz - 1 ? 0.5f * qo[c] * w.wt : half_up(w.wt, qo[c], qo[c + sz]);  // :169
    return f;
}

// Both half steps of upwind_flux3 (adv_fluxcorr.f90:158-185): the first from
// qo, the second added from dumb_q.
__device__ __forceinline__ Six upwind_total_six(const CellWinds &w, const float *__restrict__ qo,
                                                const float *__restrict__ dq, const Cell &p,
                                                const Dims &d)
{
    const off c = p.c, sj = (off)d.nx * d.nz, sz = d.nx;
    Six f = upwind_half_six(w, qo, p, d);
    f.xw = f.xw + half_up(w.uw, dq[c - 1], dq[c]);
    f.xe = f.xe + half_up(w.ue, dq[c], dq[c + 1]);
    f.ys = f.ys + half_up(w.vs, dq[c - sj], dq[c]);
    f.yn = f.yn + half_up(w.vn, dq[c], dq[c + sj]);
    if (p.k > 0) f.zb = f.zb + half_up(w.wb, dq[c - sz], dq[c]);
    f.zt = p.k == d.nz - 1 ? f.zt + 0.5f * dq[c] * w.wt : f.zt + half_up(w.wt, dq[c], dq[c + sz]);
    return f;
}

__device__ __forceinline__ bool in_tile(const StdGeom &g, const Cell &p)
{
    return p.i >= g.its && p.i <= g.ite && p.j >= g.jts && p.j <= g.jte;
}

__device__ __forceinline__ bool in_ext(const StdGeom &g, const Cell &p)
{
    return p.i >= g.its - 1 && p.i <= g.ite + 1 && p.j >= g.jts - 1 && p.j <= g.jte + 1;
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
// one RK3 stage / Euler step without the limiter (MODE 0), the same plus the
// dumb_q half step (MODE 1), or dumb_q alone (MODE 2)
// ---------------------------------------------------------------------------
template <int H, int V, int MODE>
__global__ void __launch_bounds__(256) k_std_stage(StdGeom g, float t, StdBatch b)
{
    const Dims d = g.d;
    const Cell p = this_cell(d);
    if (p.i >= d.nx || p.k >= d.nz) return;
    const bool tile = in_tile(g, p);
    const bool ext = MODE >= 1 && in_ext(g, p);
    CellWinds w{};
    if (tile || ext) w = load_winds(g, p);
    for (int s = 0; s < b.n; s++) {
        const float *__restrict__ qo = b.qold[s];
        const float q0 = qo[p.c];
        if (MODE <= 1) {
            float r = q0;  // qfluxes = qold outside the tile (:298)
            if (tile) r = flux_update(q0, flux3_six<H, V>(w, t, b.qin[s], p, d), w);
            b.out[s][p.c] = r;
        }
        if (MODE >= 1) {
            float r = q0;  // dumb_q = q outside the extended tile (:171)
            if (ext) r = flux_update(q0, upwind_half_six(w, qo, p, d), w);
            b.dq[s][p.c] = r;
        }
    }
}

// ---------------------------------------------------------------------------
// WRF_flux_corr up to scale_in / scale_out, per extended-tile cell
// ---------------------------------------------------------------------------
template <int H, int V>
__global__ void __launch_bounds__(256) k_std_scale(StdGeom g, StdBatch b)
{
    const Dims d = g.d;
    const Cell p = this_cell(d);
    if (p.i >= d.nx || p.k >= d.nz || !in_ext(g, p)) return;
    const CellWinds w = load_winds(g, p);
    const off c = p.c, sj = (off)d.nx * d.nz, sz = d.nx;
    for (int s = 0; s < b.n; s++) {
        const float *__restrict__ qo = b.qold[s];
        const Six uf = upwind_total_six(w, qo, b.dq[s], p, d);
        const float q0 = qo[c];
        // :75-78 the upwind-only value
        const float temp = flux_update(q0, uf, w);
        // :80-85 merged fluxes (flux3 with t_factor 1)
        const Six hf = flux3_six<H, V>(w, 1.0f, b.qin[s], p, d);
        const float mxw = hf.xw - uf.xw, mxe = hf.xe - uf.xe, mys = hf.ys - uf.ys,
                    myn = hf.yn - uf.yn, mzb = hf.zb - uf.zb, mzt = hf.zt - uf.zt;
        // :88-96
        float fin = -((fmn(0.0f, mxe) - fmx(0.0f, mxw)) + (fmn(0.0f, myn) - fmx(0.0f, mys))) / w.rj;
        fin = fin - (fmn(0.0f, mzt) - fmx(0.0f, mzb)) / w.rdj;
        float fout = ((fmx(0.0f, mxe) - fmn(0.0f, mxw)) + (fmx(0.0f, myn) - fmn(0.0f, mys))) / w.rj;
        fout = fout + (fmx(0.0f, mzt) - fmn(0.0f, mzb)) / w.rdj;
        // :99-121 bounds from the upwind neighbours (unscaled winds)
        float qmax = q0, qmin = q0;
        if (w.uw > 0.0f) {
            qmax = fmx(qo[c - 1], qmax);
            qmin = fmn(qo[c - 1], qmin);
        } else if (w.ue < 0.0f) {
            qmax = fmx(qo[c + 1], qmax);
            qmin = fmn(qo[c + 1], qmin);
        }
        if (w.vs > 0.0f) {
            qmax = fmx(qo[c - sj], qmax);
            qmin = fmn(qo[c - sj], qmin);
        } else if (w.vn < 0.0f) {
            qmax = fmx(qo[c + sj], qmax);
            qmin = fmn(qo[c + sj], qmin);
        }
        if (p.k > 0 && w.wb > 0.0f) {
            qmax = fmx(qo[c - sz], qmax);
            qmin = fmn(qo[c - sz], qmin);
        }
        if (p.k < d.nz - 1 && w.wt < 0.0f) {
            qmax = fmx(qo[c + sz], qmax);
            qmin = fmn(qo[c + sz], qmin);
        }
        // :124-125
        float si = 1.0f, so = 1.0f;
        if (fin > (qmax - temp)) si = fmx(0.0f, (qmax - temp) / (fin + 1.E-15f));
        if (fout > (temp - qmin)) so = fmx(0.0f, (temp - qmin) / (fout + 1.E-15f));
        b.sin[s][c] = si;
        b.sout[s][c] = so;
    }
}

// a merged flux m through a face whose positive-side cell is cp and
// negative-side cell cn, normalised as adv_fluxcorr.f90:129-145 does
__device__ __forceinline__ float limit_face(float m, const float *__restrict__ si,
                                            const float *__restrict__ so, off cp, off cn)
{
    if (m > 0.0f) return fmn(si[cp], so[cn]) * m;
    if (m < 0.0f) return fmn(so[cp], si[cn]) * m;
    return m;
}

// ---------------------------------------------------------------------------
// the limited fluxes and the final update, per tile cell
// ---------------------------------------------------------------------------
template <int H, int V>
__global__ void __launch_bounds__(256) k_std_final(StdGeom g, StdBatch b)
{
    const Dims d = g.d;
    const Cell p = this_cell(d);
    if (p.i >= d.nx || p.k >= d.nz) return;
    const bool tile = in_tile(g, p);
    CellWinds w{};
    if (tile) w = load_winds(g, p);
    const off c = p.c, sj = (off)d.nx * d.nz, sz = d.nx;
    for (int s = 0; s < b.n; s++) {
        const float *__restrict__ qo = b.qold[s];
        const float q0 = qo[c];
        if (!tile) {
            b.out[s][c] = q0;
            continue;
        }
        const float *__restrict__ si = b.sin[s];
        const float *__restrict__ so = b.sout[s];
        const Six uf = upwind_total_six(w, qo, b.dq[s], p, d);
        const Six hf = flux3_six<H, V>(w, 1.0f, b.qin[s], p, d);
        Six f;
        f.xw = limit_face(hf.xw - uf.xw, si, so, c, c - 1) + uf.xw;
        f.xe = limit_face(hf.xe - uf.xe, si, so, c + 1, c) + uf.xe;
        f.ys = limit_face(hf.ys - uf.ys, si, so, c, c - sj) + uf.ys;
        f.yn = limit_face(hf.yn - uf.yn, si, so, c + sj, c) + uf.yn;
        // the ground face stays 0 - 0 + 0; the model top is merged but not limited
        f.zb = p.k == 0 ? (hf.zb - uf.zb) + uf.zb : limit_face(hf.zb - uf.zb, si, so, c, c - sz) + uf.zb;
        f.zt = p.k == d.nz - 1 ? (hf.zt - uf.zt) + uf.zt
                               : limit_face(hf.zt - uf.zt, si, so, c + sz, c) + uf.zt;
        b.out[s][c] = flux_update(q0, f, w);
    }
}

dim3 std_grid(const Dims &d) { return dim3((d.nx + 63) / 64, (d.nz + 3) / 4, d.ny); }

template <int H, int V>
void launch_std_orders(hipStream_t s, int what, const StdGeom &g, float t, int mode,
                       const StdBatch &b)
{
    const dim3 grid = std_grid(g.d), block(64, 4);
    if (what == 0) {
        if (mode == 0) hipLaunchKernelGGL((k_std_stage<H, V, 0>), grid, block, 0, s, g, t, b);
        if (mode == 1) hipLaunchKernelGGL((k_std_stage<H, V, 1>), grid, block, 0, s, g, t, b);
        if (mode == 2) hipLaunchKernelGGL((k_std_stage<H, V, 2>), grid, block, 0, s, g, t, b);
    } else if (what == 1) {
        hipLaunchKernelGGL((k_std_scale<H, V>), grid, block, 0, s, g, b);
    } else {
        hipLaunchKernelGGL((k_std_final<H, V>), grid, block, 0, s, g, b);
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
    hipLaunchKernelGGL(k_std_wind, std_grid(d), dim3(64, 4), 0, s, d, dx, dt, adv_den, f, Um, Vm,
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

}  // namespace icaradv
