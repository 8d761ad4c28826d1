// Fused, species-batched MPDATA path (kernel_variant 1).
//
// One advect() step of mpdata_order n (adv_mpdata.f90:352-413) becomes
//   k_coef     species-invariant face coefficients of the pseudo-velocities
//              (|U|(1-|U|/avg G), the cross-term wind products and the face
//              sums of G), once per step
//   k_first    the donor-cell pass (upwind_advection, :43-103) for every
//              species in one launch, the Courant numbers loaded once per
//              cell and reused across species (bit-identical to the
//              reference order)
//   k_corr     one launch per corrective pass (iord >= 2) for every species:
//              antidiffusive pseudo-velocities (mpdata_fluxes, :105-252),
//              the 1-D FCT limiter in x, y and z (flux_limiter, :254-350 and
//              adv_mpdata_FCT_core.f90) and the donor-cell flux divergence
//              with the limited velocities, fused in one kernel.
// k_courant_b mpdata_compute_wind for every domain of the batch
// k_corr marches upward through the levels of a 60 x 16 column tile (halo 2,
// 64 x 20 columns staged in LDS); x/y neighbours come from LDS planes, the z
// direction (including the column-wise FCT) stays in registers of the thread
// that owns the column.  Species are independent inside advect, so a launch
// covers (tile, species) pairs; blocks are mapped so that the species of one
// tile run on the same XCD and share its L2.
//
// Every expression keeps the reference's evaluation order, with IEEE FP32
// division and no contraction (this unit is compiled with
// -ffp-contract=off), so for FP32 storage the result is bit-identical to the
// exact kernels and to a strict IEEE FP32 build of the reference; the speed
// comes from fusion, batching and the per-step invariants, not from cheaper
// arithmetic.  The MPDATA step on rough fields is sensitive to rounding
// order at the 1e-5..1e-3 level (a reordered build of the reference moves
// theta by ~6e-4 in 5 steps), so no reordering is done.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <type_traits>

#include "../../include/icar_adv.h"
#include "fast.hpp"
#include "fdiv.hpp"

namespace icaradv {

namespace {

// ---------------------------------------------------------------------------
// Pointers read from the descriptor tables in device memory are generic to
// the compiler, which would then emit flat loads (64-bit per-lane addresses,
// both memory counters).  Every buffer here is device memory, so the kernels
// view them in the global address space: global loads with a scalar base
// and a 32-bit lane offset.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) const float gcfloat;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const uint16_t gcu16;
typedef __attribute__((address_space(1))) uint16_t gu16;

__device__ __forceinline__ gcfloat *gp(const float *p) { return (gcfloat *)p; }
__device__ __forceinline__ gfloat *gp(float *p) { return (gfloat *)p; }

// ---------------------------------------------------------------------------
// species storage: FP32 or 16-bit (BF16 / FP16) holding value * 2^-e
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool is_bf16(int kind)
{
    return kind == ICAR_ADV_STORE_BF16 || kind == ICAR_ADV_STORE_BF16_SR;
}

// A species value in two steps: the raw bits (a 32- or 16-bit load), then the
// conversion.  Kernels that keep several loads in flight load raw bits first
// and convert when the values are used, so that no conversion right after a
// load makes the wave wait for it.
typedef __attribute__((address_space(1))) const uint32_t gcu32;

// One 32-bit load for either width: FP32 element i, or the aligned pair of
// 16-bit elements holding element i (16-bit buffers are padded so that the
// pair of the last element is in bounds).  The same instruction for every
// storage kind, so a batch mixing kinds needs no branch around its loads.
__device__ __forceinline__ uint32_t ld_raw(const void *p, size_t i, int kind)
{
    return ((gcu32 *)p)[kind == ICAR_ADV_STORE_F32 ? i : (i >> 1)];
}

// the value of element i from the word ld_raw returned for it, with a branch
// on the kind: fewer live registers where the kind is the same for the block
__device__ __forceinline__ float cvt_q_uniform(uint32_t w, size_t i, int kind, float scale)
{
    if (kind == ICAR_ADV_STORE_F32) return __uint_as_float(w);
    const uint32_t raw = (i & 1) ? (w >> 16) : (w & 0xFFFFu);
    const float v = is_bf16(kind) ? __uint_as_float(raw << 16)
                                  : __half2float(__ushort_as_half((uint16_t)raw));
    return v * scale;
}

// the same without a branch: both 16-bit forms are decoded and the kind
// selects, so that no conversion waits on a branch while loads are in flight
__device__ __forceinline__ float cvt_q(uint32_t w, size_t i, int kind, float scale)
{
    const uint32_t raw = (i & 1) ? (w >> 16) : (w & 0xFFFFu);
    const float vb = __uint_as_float(raw << 16);
    const float vh = __half2float(__ushort_as_half((uint16_t)raw));
    const float v16 = (is_bf16(kind) ? vb : vh) * scale;
    return kind == ICAR_ADV_STORE_F32 ? __uint_as_float(w) : v16;
}

__device__ __forceinline__ float ld_q(const void *p, size_t i, int kind, float scale)
{
    if (kind == ICAR_ADV_STORE_F32) return ((gcfloat *)p)[i];
    const uint16_t b = ((gcu16 *)p)[i];
    const float v = is_bf16(kind) ? __uint_as_float((uint32_t)b << 16)
                                  : __half2float(__ushort_as_half(b));
    return v * scale;
}

// 16 random bits for the stochastic rounding of cell c of species s at step t:
// a counter-based hash (the murmur3 32-bit finaliser of a mixed key), so the
// rounding is deterministic and reproducible off the device.
__device__ __forceinline__ uint32_t sr_bits(uint32_t c, int s, uint32_t t)
{
    uint32_t x = (c * 0x9E3779B1u) ^ ((uint32_t)(s + 1) * 0x85EBCA77u) ^ (t * 0xC2B2AE3Du);
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x & 0xFFFFu;
}

// BF16 bits of x: stochastic rounding with the 16 random bits rnd (sr), else
// round to nearest even; infinities and NaN are kept (a NaN stays a NaN)
__device__ __forceinline__ uint16_t bf16_of(float x, bool sr, uint32_t rnd)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += sr ? rnd : 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// value -> 16-bit storage of value * inv_scale.  The plain kinds round to
// nearest even; the _SR kinds round up with probability equal to the
// discarded fraction of an ulp (rnd: 16 uniform bits), so that increments
// smaller than half an ulp are not lost on average and the flux form keeps
// conserving mass.
__device__ __forceinline__ uint16_t to16(float v, int kind, float inv_scale, uint32_t rnd)
{
    const float x = v * inv_scale;
    if (is_bf16(kind)) return bf16_of(x, kind == ICAR_ADV_STORE_BF16_SR, rnd);
    if (kind == ICAR_ADV_STORE_F16_SR) {
        const float ax = fabsf(x);
        if (ax >= 6.103515625e-05f) {  // FP16 normal range: 13 mantissa bits are dropped
            const uint32_t u = (__float_as_uint(x) + (rnd & 0x1FFFu)) & ~0x1FFFu;
            return __half_as_ushort(__float2half_rz(__uint_as_float(u)));
        }
        // FP16 subnormals are multiples of 2^-24
        const float a = ax * 16777216.0f;
        const float k = floorf(a);
        const float up = ((float)rnd * 1.52587890625e-05f < a - k) ? 1.0f : 0.0f;
        return __half_as_ushort(__float2half_rn(copysignf((k + up) * 5.9604644775390625e-08f, x)));
    }
    return __half_as_ushort(__float2half_rn(x));
}

__device__ __forceinline__ void st_q(void *p, size_t i, int kind, float inv_scale, float v,
                                     uint32_t rnd)
{
    if (kind == ICAR_ADV_STORE_F32)
        ((gfloat *)p)[i] = v;
    else
        ((gu16 *)p)[i] = to16(v, kind, inv_scale, rnd);
}

__device__ __forceinline__ bool is_sr(int kind)
{
    return kind == ICAR_ADV_STORE_BF16_SR || kind == ICAR_ADV_STORE_F16_SR;
}

// The storage kinds a kernel is compiled for: FP32 only; FP32 and the BF16
// kinds; any kind.  A species' kind is the same for a whole block (or, in the
// donor-cell kernel, for a species of the loop), so the BF16 class branches on
// it around its loads: one 16-bit load of a BF16 element (zero-extended) or
// one 32-bit load of an FP32 one.  The branch holds the load only, so that the
// load stays in flight past it; the conversion (a shift into the high half
// and the scale) comes where the value is used.  No FP16 path.
constexpr int QK_F32 = 0, QK_BF16 = 1, QK_ANY = 2;

// the raw word of element i: the float's bits (FP32), the 16 bits (BF16
// class, a 16-bit kind), or the ld_raw word (any kind)
template <int QK>
__device__ __forceinline__ uint32_t ld_w(const void *p, size_t i, int kind)
{
    if (QK == QK_F32) return ((gcu32 *)p)[i];
    if (QK == QK_BF16) {
        if (kind == ICAR_ADV_STORE_F32) return ((gcu32 *)p)[i];
        return ((gcu16 *)p)[i];
    }
    return ld_raw(p, i, kind);
}

// the value of element i from its ld_w word
template <int QK>
__device__ __forceinline__ float cv_w(uint32_t w, size_t i, int kind, float scale)
{
    if (QK == QK_F32) return __uint_as_float(w);
    if (QK == QK_BF16)
        return kind == ICAR_ADV_STORE_F32 ? __uint_as_float(w) : __uint_as_float(w << 16) * scale;
    return cvt_q(w, i, kind, scale);
}

// the same where the kind is uniform for the block (any kind: a branch, not a
// select)
template <int QK>
__device__ __forceinline__ float cv_w_uniform(uint32_t w, size_t i, int kind, float scale)
{
    if (QK == QK_ANY) return cvt_q_uniform(w, i, kind, scale);
    return cv_w<QK>(w, i, kind, scale);
}

template <int QK>
__device__ __forceinline__ void st_w(void *p, size_t i, int kind, float inv_scale, float v,
                                     uint32_t rnd)
{
    if (QK == QK_F32) {
        ((gfloat *)p)[i] = v;
    } else if (QK == QK_BF16) {
        if (kind == ICAR_ADV_STORE_F32)
            ((gfloat *)p)[i] = v;
        else
            ((gu16 *)p)[i] = bf16_of(v * inv_scale, kind == ICAR_ADV_STORE_BF16_SR, rnd);
    } else {
        st_q(p, i, kind, inv_scale, v, rnd);
    }
}

__device__ __forceinline__ float max3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }
__device__ __forceinline__ float min3(float a, float b, float c) { return fminf(fminf(a, b), c); }

// One pass of one species: where to read q0 (and q1) and where to write.
struct SpecPass {
    const void *q0;   // field before the pass (storage kind0)
    const float *q1;  // field after the previous pass (FP32), or nullptr if equal to q0
    void *out;        // result (storage kindo)
    int kind0, kindo;
    float scale0, inv_scaleo;
};

// One domain of a batched launch (a single handle, or one member of an
// ensemble of unevenly sized domains).  Cell kernels cover 64x4-thread
// blocks of one level; the corrective kernel covers (tile, species) items.
constexpr int MAXP = 16;  // mpdata_order limit
struct DomDesc {
    Dims d;
    unsigned cb0, bx;     // first global cell block; cell blocks along x
    unsigned w0;          // first global corrective work item
    int tiles_x, nspec;
    int nks, kch;         // corrective kernel: column chunks along k and levels per chunk
    float dx, dt;
    int advect_density;
    int compensated;      // flux divergence with compensated sums
    DomainFields f;       // raw inputs
    StepFields sf;        // Courant numbers and products (written by k_courant_b)
    float *coef;          // face coefficients (float4 groups)
    unsigned *step;       // steps taken (drives the stochastic rounding)
    const SpecPass *pass[MAXP];  // species table of pass p at index p-1
};

// domain of global cell block b / work item w: the last entry whose first
// block (item) is <= the key
__device__ __forceinline__ int find_dom_cb(const DomDesc *D, int nd, unsigned b)
{
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (D[mid].cb0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_dom_w(const DomDesc *D, int nd, unsigned w)
{
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (D[mid].w0 <= w) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// (i, k, j) of this thread in a 64x4 cell block
__device__ __forceinline__ void cell_of(const DomDesc &D, unsigned b, int &i, int &k, int &j)
{
    const unsigned l = b - D.cb0;
    const unsigned r = l / D.bx;
    i = (int)(l % D.bx) * 64 + (int)threadIdx.x;
    k = (int)(r % (unsigned)D.d.nz);
    j = (int)(r / (unsigned)D.d.nz) * 4 + (int)threadIdx.y;
}

// ---------------------------------------------------------------------------
// mpdata_compute_wind (adv_mpdata.f90:477-504) and the species-invariant
// products, for every domain of the batch; the same expressions as the
// exact Courant kernel, so the same bits.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_courant_b(const DomDesc *__restrict__ DD, int nd)
{
    const int dm = find_dom_cb(DD, nd, blockIdx.x);
    const DomDesc &D = DD[dm];
    int i, k, j;
    cell_of(D, blockIdx.x, i, k, j);
    const Dims d = D.d;
    if (i >= d.nx || j >= d.ny) return;
    gcfloat *__restrict__ den = gp(D.f.density), *__restrict__ fu = gp(D.f.u);
    gcfloat *__restrict__ fv = gp(D.f.v), *__restrict__ fw = gp(D.f.w);
    gcfloat *__restrict__ ju = gp(D.f.jaco_u), *__restrict__ jv = gp(D.f.jaco_v);
    gcfloat *__restrict__ jw = gp(D.f.jaco_w), *__restrict__ jc = gp(D.f.jaco);
    gcfloat *__restrict__ dz = gp(D.f.dz);
    gfloat *__restrict__ oU = gp(D.sf.U), *__restrict__ oV = gp(D.sf.V), *__restrict__ oW = gp(D.sf.W);
    gfloat *__restrict__ oG = gp(D.sf.G), *__restrict__ oGz = gp(D.sf.Gz), *__restrict__ oWd = gp(D.sf.Wd);
    const float dt = D.dt, dx = D.dx;
    const int advect_density = D.advect_density;
    const size_t c = ix(d, i, k, j);
    const size_t sj = (size_t)d.nx * d.nz;
    const float rc = advect_density ? den[c] : 1.0f;  // :489-490
    if (i >= 1) {
        const size_t iu = (size_t)i + (size_t)(d.nx + 1) * ((size_t)k + (size_t)d.nz * j);
        const float rl = advect_density ? den[c - 1] : 1.0f;
        oU[c] = fu[iu] * dt * (rc + rl) * 0.5f * ju[iu] / dx;  // :492-493
    } else {
        oU[c] = 0.0f;
    }
    if (j >= 1) {
        const float rs = advect_density ? den[c - sj] : 1.0f;
        oV[c] = fv[c] * dt * (rc + rs) * 0.5f * jv[c] / dx;  // :494-495
    } else {
        oV[c] = 0.0f;
    }
    float Wc;
    if (k < d.nz - 1) {
        const float ru = advect_density ? den[c + d.nx] : 1.0f;
        Wc = fw[c] * dt * jw[c] * (ru + rc) * 0.5f;  // :496-497
    } else {
        Wc = fw[c] * dt * jw[c] * rc;  // :498
    }
    oW[c] = Wc;
    const float jac = jc[c], dnc = den[c], dzc = dz[c];
    oG[c] = jac * dnc;
    oGz[c] = dzc * jac * dnc;
    oWd[c] = Wc / dzc;
}

// ---------------------------------------------------------------------------
// species-invariant parts of the pseudo-velocities, once per step, each
// evaluated exactly as mpdata_fluxes evaluates it (adv_mpdata.f90:139,151,
// 163,180,192,203,221,233,244):
//   CUx = |U|*(1-|U|/(0.5*(G_r+G_l)))   first factor of the face value
//   CVx = (0.5*U)*edge_v                 prefix of each cross term
//   SGx = G_r+G_l                        its divisor
// for the x face left of a cell (x), the y face south of it (y) and the z
// face on top of it (z).  A cross-term prefix is 0 where the reference skips
// the term (its row/level guards); the corrective kernel tests the same
// guards and never uses those entries.  Stored as three float4 groups per
// cell (x, y, z faces) so that the corrective kernel loads a face's four
// values with one 128-bit load.
// ---------------------------------------------------------------------------
enum { CUX = 0, CXV, CXW, SGX, CUY, CYU, CYW, SGY, CUZ, CZU, CZV, SGZ, NCOEF };

__global__ void __launch_bounds__(256) k_coef(const DomDesc *__restrict__ DD, int nd)
{
    const int dm = find_dom_cb(DD, nd, blockIdx.x);
    const DomDesc &D = DD[dm];
    int i, k, j;
    cell_of(D, blockIdx.x, i, k, j);
    const Dims d = D.d;
    if (i >= d.nx || j >= d.ny) return;
    gcfloat *__restrict__ U = gp(D.sf.U), *__restrict__ V = gp(D.sf.V), *__restrict__ Wd = gp(D.sf.Wd);
    gcfloat *__restrict__ G = gp(D.sf.G);
    gfloat *__restrict__ coef = gp(D.coef);
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const size_t N = (size_t)nx * ny * nz;
    const size_t sk = (size_t)nx, sj = (size_t)nx * nz;
    const size_t c = (size_t)i + sk * ((size_t)k + (size_t)nz * j);
    float v[NCOEF];
#pragma unroll
    for (int e = 0; e < NCOEF; e++) v[e] = 0.0f;
    if (i >= 1) {
        const float Uc = U[c], au = fabsf(Uc);
        v[CUX] = au * (1.0f - au / (0.5f * (G[c] + G[c - 1])));
        v[SGX] = G[c] + G[c - 1];
        if (j >= 1 && j <= ny - 2)
            v[CXV] = 0.5f * Uc * ((1.0f / 4.0f) * (V[c] + V[c + sj] + V[c - 1] + V[c - 1 + sj]));
        if (k >= 1 && k <= nz - 2)
            v[CXW] = 0.5f * Uc * ((1.0f / 4.0f) * (Wd[c] + Wd[c - sk] + Wd[c - 1] + Wd[c - 1 - sk]));
    }
    if (j >= 1) {
        const float Vc = V[c], av = fabsf(Vc);
        v[CUY] = av * (1.0f - av / (0.5f * (G[c] + G[c - sj])));
        v[SGY] = G[c] + G[c - sj];
        if (i >= 1 && i <= nx - 2)
            v[CYU] = 0.5f * Vc * ((1.0f / 4.0f) * (U[c + 1] + U[c + 1 - sj] + U[c] + U[c - sj]));
        if (k >= 1 && k <= nz - 2)
            v[CYW] = 0.5f * Vc * ((1.0f / 4.0f) * (Wd[c] + Wd[c - sk] + Wd[c - sj] + Wd[c - sk - sj]));
    }
    if (k <= nz - 2) {
        const float Wc = Wd[c], aw = fabsf(Wc);
        v[CUZ] = aw * (1.0f - aw / (0.5f * (G[c + sk] + G[c])));
        v[SGZ] = G[c + sk] + G[c];
        if (i >= 1 && i <= nx - 2)
            v[CZU] = 0.5f * Wc * ((1.0f / 4.0f) * (U[c + 1] + U[c + 1 + sk] + U[c] + U[c + sk]));
        if (j >= 1 && j <= ny - 2)
            v[CZV] = 0.5f * Wc * ((1.0f / 4.0f) * (V[c] + V[c + sk] + V[c + sj] + V[c + sk + sj]));
    }
#pragma unroll
    for (int g = 0; g < 3; g++)
#pragma unroll
        for (int e = 0; e < 4; e++) coef[4 * ((size_t)g * N + c) + e] = v[4 * g + e];
}

// one step taken by every domain of the batch (a vector store per domain)
__global__ void __launch_bounds__(64) k_tick(const DomDesc *__restrict__ DD, int nd)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t < nd) DD[t].step[0] = DD[t].step[0] + 1u;
}

// ---------------------------------------------------------------------------
// donor-cell pass for every species (upwind_advection, adv_mpdata.f90:43-103)
// ---------------------------------------------------------------------------
constexpr int FBX = 64, FBY = 4;
constexpr int FSB = 4;  // species per load group in k_first

// flux1 (adv_mpdata.f90:39)
__device__ __forceinline__ float flux1(float l, float r, float U)
{
    return ((U + fabsf(U)) * l + (U - fabsf(U)) * r) / 2.0f;
}

// FD: the two divisions of each species by G and Gz (the same for every
// species of the cell: one reciprocal each) as the three-operation quotient
// of fdiv.hpp where the operands are in range, else, for the whole wave, the
// IEEE division
template <int QK, bool FD>
__global__ void __launch_bounds__(FBX *FBY) k_first(const DomDesc *__restrict__ DD, int nd)
{
    const int dm = find_dom_cb(DD, nd, blockIdx.x);
    const DomDesc &D = DD[dm];
    int i, k, j;
    cell_of(D, blockIdx.x, i, k, j);
    const Dims d = D.d;
    if (i >= d.nx || j >= d.ny) return;
    gcfloat *__restrict__ U = gp(D.sf.U), *__restrict__ V = gp(D.sf.V), *__restrict__ W = gp(D.sf.W);
    gcfloat *__restrict__ G = gp(D.sf.G), *__restrict__ Gz = gp(D.sf.Gz);
    const SpecPass *__restrict__ sp = D.pass[0];
    const int nspec = D.nspec;
    const size_t sk = (size_t)d.nx, sj = (size_t)d.nx * d.nz;
    const size_t c = (size_t)i + sk * ((size_t)k + (size_t)d.nz * j);
    const bool ring = i == 0 || i == d.nx - 1 || j == 0 || j == d.ny - 1;
    float Ul = 0, Ur = 0, Vs = 0, Vn = 0, Wt = 0, Wb = 0, g = 1, gz = 1;
    if (!ring) {
        Ul = U[c];
        Ur = U[c + 1];
        Vs = V[c];
        Vn = V[c + sj];
        Wt = W[c];
        Wb = k > 0 ? W[c - sk] : 0.0f;
        g = G[c];
        gz = Gz[c];
    }
    const unsigned tstep = *D.step;
    const float yg = (FD && !ring) ? recip_nr(g) : 1.0f, ygz = (FD && !ring) ? recip_nr(gz) : 1.0f;
    // Species in groups of FSB: every load of a group is issued before the
    // first store, so a thread keeps several species' reads in flight (the
    // compiler cannot move loads above the stores of other species itself:
    // the table's pointers may alias).
    for (int s0 = 0; s0 < nspec; s0 += FSB) {
        uint32_t rc[FSB], rw[FSB], re[FSB], rs[FSB], rn[FSB], ru[FSB], rd[FSB];
#pragma unroll
        for (int b = 0; b < FSB; b++) {
            rc[b] = rw[b] = re[b] = rs[b] = rn[b] = ru[b] = rd[b] = 0u;
            if (s0 + b < nspec) {
                SpecPass p = sp[s0 + b];
                if (QK == QK_F32) p.kind0 = ICAR_ADV_STORE_F32;  // an FP32-only batch: no 16-bit code
                p.kind0 = __builtin_amdgcn_readfirstlane(p.kind0);
                rc[b] = ld_w<QK>(p.q0, c, p.kind0);
                if (!ring) {
                    rw[b] = ld_w<QK>(p.q0, c - 1, p.kind0);
                    re[b] = ld_w<QK>(p.q0, c + 1, p.kind0);
                    rs[b] = ld_w<QK>(p.q0, c - sj, p.kind0);
                    rn[b] = ld_w<QK>(p.q0, c + sj, p.kind0);
                    if (k < d.nz - 1) ru[b] = ld_w<QK>(p.q0, c + sk, p.kind0);
                    if (k > 0) rd[b] = ld_w<QK>(p.q0, c - sk, p.kind0);
                }
            }
        }
        float qc[FSB], qw[FSB], qe[FSB], qs[FSB], qn[FSB], qu[FSB], qd[FSB];
#pragma unroll
        for (int b = 0; b < FSB; b++) {
            int kind = ICAR_ADV_STORE_F32;
            float sc = 1.0f;
            if (QK != QK_F32 && s0 + b < nspec) {
                kind = __builtin_amdgcn_readfirstlane(sp[s0 + b].kind0);
                sc = sp[s0 + b].scale0;
            }
            if (QK == QK_BF16) {
                // one branch per species (the kind is uniform): FP32 words as
                // they are, BF16 ones shifted into place and scaled
                if (kind == ICAR_ADV_STORE_F32) {
                    qc[b] = __uint_as_float(rc[b]);
                    qw[b] = __uint_as_float(rw[b]);
                    qe[b] = __uint_as_float(re[b]);
                    qs[b] = __uint_as_float(rs[b]);
                    qn[b] = __uint_as_float(rn[b]);
                    qu[b] = __uint_as_float(ru[b]);
                    qd[b] = __uint_as_float(rd[b]);
                } else {
                    qc[b] = __uint_as_float(rc[b] << 16) * sc;
                    qw[b] = __uint_as_float(rw[b] << 16) * sc;
                    qe[b] = __uint_as_float(re[b] << 16) * sc;
                    qs[b] = __uint_as_float(rs[b] << 16) * sc;
                    qn[b] = __uint_as_float(rn[b] << 16) * sc;
                    qu[b] = __uint_as_float(ru[b] << 16) * sc;
                    qd[b] = __uint_as_float(rd[b] << 16) * sc;
                }
                continue;
            }
            qc[b] = cv_w<QK>(rc[b], c, kind, sc);
            qw[b] = cv_w<QK>(rw[b], c - 1, kind, sc);
            qe[b] = cv_w<QK>(re[b], c + 1, kind, sc);
            qs[b] = cv_w<QK>(rs[b], c - sj, kind, sc);
            qn[b] = cv_w<QK>(rn[b], c + sj, kind, sc);
            qu[b] = cv_w<QK>(ru[b], c + sk, kind, sc);
            qd[b] = cv_w<QK>(rd[b], c - sk, kind, sc);
        }
        float qo[FSB];
        auto update = [&](auto fd_c) -> bool {
            constexpr bool F = decltype(fd_c)::value;
            DivGuard gd;
            if (F) {
                guard_den(gd, g);
                guard_den(gd, gz);
            }
#pragma unroll
            for (int b = 0; b < FSB; b++) {
                qo[b] = qc[b];
                if (s0 + b >= nspec || ring) continue;  // the ring is copied (:61-63)
                const float fr = flux1(qc[b], qe[b], Ur), fl = flux1(qw[b], qc[b], Ul);
                const float fn = flux1(qc[b], qn[b], Vn), fs = flux1(qs[b], qc[b], Vs);
                const float xh = (fr - fl) + (fn - fs);  // :86-87
                float xv;
                if (k == 0)
                    xv = flux1(qc[b], qu[b], Wt);  // :93-94
                else if (k == d.nz - 1)
                    xv = qc[b] * Wt - flux1(qd[b], qc[b], Wb);  // top flux not upwinded (:96-97)
                else
                    xv = flux1(qc[b], qu[b], Wt) - flux1(qd[b], qc[b], Wb);  // :90-91
                if constexpr (F) {
                    guard_num(gd, xh);
                    guard_num(gd, xv);
                    const float q = qc[b] - div_y(xh, g, yg);
                    qo[b] = q - div_y(xv, gz, ygz);
                } else {
                    const float q = qc[b] - xh / g;
                    qo[b] = q - xv / gz;
                }
            }
            return guard_ok(gd);
        };
        if (!FD || __any(!update(std::integral_constant<bool, true>{})))
            update(std::integral_constant<bool, false>{});
#pragma unroll
        for (int b = 0; b < FSB; b++) {
            const int s = s0 + b;
            if (s >= nspec) continue;
            SpecPass p = sp[s];
            if (QK == QK_F32) p.kindo = ICAR_ADV_STORE_F32;
            p.kindo = __builtin_amdgcn_readfirstlane(p.kindo);
            const uint32_t rnd = is_sr(p.kindo) ? sr_bits((uint32_t)c, s, tstep) : 0u;
            st_w<QK>(p.out, c, p.kindo, p.inv_scaleo, qo[b], rnd);
        }
    }
}

// ---------------------------------------------------------------------------
// fused corrective pass
// ---------------------------------------------------------------------------
constexpr int CX = 64;            // staged tile width (one wavefront per row)
constexpr int CH = 2;             // halo
constexpr int CTI = CX - 2 * CH;  // 60 output columns
constexpr int OWN = 2;            // output rows per wave: ly + NW*(rb + t), t = 0, 1

// A corrective tile of TJ output rows (16 or 32): CY = TJ + 4 staged rows,
// NW = TJ / 2 wavefronts (wave ly stages rows ly, ly + NW, ... and owns two
// output rows), SR staged-row slots per wave.  Taller tiles recompute fewer
// halo rows (y faces on CY - 1 rows, y betas on CY - 2) per output row but
// take twice the LDS (one block of 16 waves per CU instead of two of 8).
template <int TJ>
struct Tile {
    static constexpr int CY = TJ + 2 * CH;
    static constexpr int NW = TJ / 2;
    static constexpr int CNT = 64 * NW;
    static constexpr int SR = (CY + NW - 1) / NW;
};


// s = RN(a + b) and its rounding error e, a + b = s + e exactly (Knuth's
// TwoSum; the unit is compiled without contraction)
__device__ __forceinline__ float two_sum(float a, float b, float &e)
{
    const float s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
    return s;
}

// The opt-in compensated form of the update q1 - ((fr-fl) + (fn-fs))/G -
// (Fzt-Fzb)/Gz (upwind_advection, adv_mpdata.f90:86-97): the flux sums and
// the two subtractions carry their rounding errors (TwoSum), the quotients
// their remainders (one FMA each), and the errors enter once, before the
// final rounding; the result is q1 minus the exact divergence of the FP32
// fluxes, rounded about once.  Not the reference's operation order.
__device__ __forceinline__ float divergence_compensated(float q1, float fl, float fr, float fs,
                                                        float fn, float fzt, float fzb, float G,
                                                        float Gz)
{
    float ea, eb, eh, ez, e1, e2;
    const float a = two_sum(fr, -fl, ea);
    const float b = two_sum(fn, -fs, eb);
    const float h = two_sum(a, b, eh);
    const float dh = h / G;
    const float ch = (__builtin_fmaf(-dh, G, h) + ((ea + eb) + eh)) / G;
    const float vz = two_sum(fzt, -fzb, ez);
    const float dz = vz / Gz;
    const float cz = (__builtin_fmaf(-dz, Gz, vz) + ez) / Gz;
    const float t = two_sum(q1, -dh, e1);
    const float u = two_sum(t, -dz, e2);
    return u + ((e1 + e2) - (ch + cz));
}

// x/d: the IEEE division, or (F) the three-operation quotient of fdiv.hpp
// with the operands added to the guard g
template <bool F>
__device__ __forceinline__ float fdiv(float x, float d, DivGuard &g)
{
    if constexpr (F) {
        guard_div(g, x, d);
        return div_y(x, d, recip_nr(d));
    } else {
        return x / d;
    }
}

// the same with y = recip_of(d, g) computed once for several numerators
template <bool F>
__device__ __forceinline__ float fdiv_y(float x, float d, float y, DivGuard &g)
{
    if constexpr (F) {
        guard_num(g, x);
        return div_y(x, d, y);
    } else {
        return x / d;
    }
}

__device__ __forceinline__ float recip_of(float d, DivGuard &g)
{
    guard_den(g, d);
    return recip_nr(d);
}

// WPE: waves per SIMD the register allocation targets (2 blocks of 8 waves
// per CU need 4); PF: load this iteration's coefficients before the first
// barrier (latency hidden by the barrier) or at their use; FD: the
// three-operation quotient where its operands are in range (fdiv.hpp).
template <bool FCT, int WPE, bool PF, int QK, bool FD, int TJ>
__global__ void __launch_bounds__(Tile<TJ>::CNT) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    k_corr(const DomDesc *__restrict__ DD, int nd, int pass, unsigned nwork)
{
    constexpr int CY = Tile<TJ>::CY, NW = Tile<TJ>::NW, SR = Tile<TJ>::SR;
    // XCD-aware mapping: consecutive work items (the species of one tile)
    // land on the same XCD (blocks are dealt round-robin over 8 XCDs) and
    // share the tile's coefficients in its L2.
    const unsigned nb = gridDim.x;
    const unsigned w = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
    if (w >= nwork) return;
    const DomDesc &D = DD[find_dom_w(DD, nd, w)];
    const unsigned lw = w - D.w0;
    const int s = (int)(lw % (unsigned)D.nspec);
    const unsigned tk = lw / (unsigned)D.nspec;
    const int kc = (int)(tk % (unsigned)D.nks);  // column chunk
    const int tile = (int)(tk / (unsigned)D.nks);
    const int tx = tile % D.tiles_x, ty = tile / D.tiles_x;
    const int nx = D.d.nx, ny = D.d.ny, nz = D.d.nz;
    // Levels [k0, k1) of the column are written by this block.  The march
    // starts three levels below k0 so that the vertical pipeline (q0 of
    // three levels, the two z faces below, the betas and flux of the level
    // below) holds the values a march from the ground would hold when it
    // reaches k0; the warm-up levels are computed but not stored.
    const int k0 = kc * D.kch;
    const int k1 = min(nz, k0 + D.kch);
    const int m0 = max(0, k0 - 3);
    // 32-bit element offsets (the host checks nx*ny*nz < 2^32) so that the
    // loads use a scalar base and one VGPR offset
    const unsigned sk = (unsigned)nx, sj = (unsigned)nx * (unsigned)nz;
    const size_t N = (size_t)nx * ny * nz;
    // ly is the same for the whole wave: as a scalar, every row offset below
    // is one too, which keeps the LDS addresses out of the vector registers
    const int lx = threadIdx.x & 63, ly = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = tx * CTI - CH + lx;
    const int j0 = ty * TJ - CH;  // global row of staged row 0
    SpecPass p = D.pass[pass - 1][s];
    if (QK == QK_F32) p.kind0 = p.kindo = ICAR_ADV_STORE_F32;  // an FP32-only batch: no 16-bit code
    p.kind0 = __builtin_amdgcn_readfirstlane(p.kind0);  // uniform: branches on the kinds are scalar
    p.kindo = __builtin_amdgcn_readfirstlane(p.kindo);
    const bool same = p.q1 == nullptr;
    const unsigned tstep = is_sr(p.kindo) ? *D.step : 0u;
    const float eps = 1e-10f, eps15 = 1e-15f;
    gcfloat *__restrict__ Gp = gp(D.sf.G), *__restrict__ Gzp = gp(D.sf.Gz), *__restrict__ dzp = gp(D.f.dz);
    gcfloat *__restrict__ q1p = gp(p.q1);
    // three groups of four coefficients per cell (x, y, z faces), one
    // 128-bit load each
    gcfloat *__restrict__ gx = gp(D.coef);
    gcfloat *__restrict__ gy = gx + 4 * N;
    gcfloat *__restrict__ gz = gx + 8 * N;

    __shared__ float sq1[4][CY][CX];  // q1 at levels m-3..m (ring by level & 3)
    __shared__ float sq0[3][CY][CX];  // q0 at levels m-2..m (ring by level % 3)
    __shared__ float sxu[CY][CX], sxf[CY][CX];  // x face left of a column: U2, flux1(U2)
    __shared__ float syu[CY][CX], syf[CY][CX];  // y face south of a row
    __shared__ float sbxi[CY][CX], sbxo[CY][CX], sbyi[CY][CX], sbyo[CY][CX];  // FCT betas

    const bool lane_in = gi >= 0 && gi < nx;
    const bool lane_out = lx >= CH && lx < CX - CH && lane_in;
    const int rb = (ly < 2) ? 1 : 0;
    const int nsr = (ly + NW * (SR - 1) < CY) ? SR : SR - 1;  // staged rows of this wave

    bool row_in[SR];
    unsigned rowbase[SR];  // gi + sj*gj (the level offset is added per use)
#pragma unroll
    for (int r = 0; r < SR; r++) {
        const int gj = j0 + ly + NW * r;
        row_in[r] = r < nsr && lane_in && gj >= 0 && gj < ny;
        rowbase[r] = row_in[r] ? (unsigned)gi + sj * (unsigned)gj : 0u;
    }
    // the owned rows rb, rb+1 with static indices (rb is 0 or 1), so that no
    // per-lane array is indexed at run time
    bool own_in[OWN];
    unsigned own_base[OWN];
#pragma unroll
    for (int t = 0; t < OWN; t++) {
        own_in[t] = rb ? row_in[t + 1] : row_in[t];
        own_base[t] = rb ? rowbase[t + 1] : rowbase[t];
    }

    // per owned column: z pipeline state
    float z0a[OWN], z0b[OWN], z0c[OWN];          // q0 at levels m-2, m-1, m
    float wu2[OWN], wu1[OWN], fz2[OWN], fz1[OWN];  // unlimited w / its flux1 on faces m-2, m-1
    float bzi[OWN], bzo[OWN];                    // betas of the cell below the current one
    float Fzp[OWN];                              // flux on the face below the output level
#pragma unroll
    for (int t = 0; t < OWN; t++) {
        z0a[t] = z0b[t] = z0c[t] = 0.0f;
        wu2[t] = wu1[t] = fz2[t] = fz1[t] = 0.0f;
        bzi[t] = bzo[t] = 1.0f;
        Fzp[t] = 0.0f;
    }

    // register prefetch of the next staged level: raw bits of q0 and the
    // FP32 q1 (unless it is q0), used only after the next barrier
    uint32_t n0r[SR];
    float n1[SR];
#pragma unroll
    for (int r = 0; r < SR; r++) {
        n0r[r] = 0u;
        n1[r] = 0.0f;
        if (row_in[r]) {
            const unsigned c = rowbase[r] + sk * (unsigned)m0;
            n0r[r] = ld_w<QK>(p.q0, c, p.kind0);
            if (!same) n1[r] = q1p[c];
        }
    }

    for (int m = m0; m <= k1 + 1; m++) {
        const int L1 = m - 1, L2 = m - 2;
        // ---------------- (A) stage level m, prefetch level m+1 and the
        //                  coefficients of this iteration
        float n0[SR];
#pragma unroll
        for (int r = 0; r < SR; r++) {
            const int row = ly + NW * r;
            n0[r] = cv_w_uniform<QK>(n0r[r], rowbase[r] + sk * (unsigned)m, p.kind0, p.scale0);
            if (m < nz && r < nsr) {
                sq0[m % 3][row][lx] = n0[r];
                sq1[m & 3][row][lx] = same ? n0[r] : n1[r];
            }
        }
#pragma unroll
        for (int t = 0; t < OWN; t++) {  // own column: q0 at levels m-2, m-1, m
            const float nv = rb ? n0[t + 1] : n0[t];
            z0a[t] = z0b[t];
            z0b[t] = z0c[t];
            z0c[t] = (m < nz) ? nv : 0.0f;
        }
        if (m + 1 < nz) {
#pragma unroll
            for (int r = 0; r < SR; r++) {
                if (row_in[r]) {
                    const unsigned c = rowbase[r] + sk * (unsigned)(m + 1);
                    n0r[r] = ld_w<QK>(p.q0, c, p.kind0);
                    if (!same) n1[r] = q1p[c];
                }
            }
        }
        // x faces and divergence at L2 (own rows), y faces at L2 (staged
        // rows), z faces at L1 (own rows)
        float kx[OWN][4], kz[OWN][5], ky[SR][4], kg[OWN][2];
        // The loads are unconditional, from a valid cell: a lane outside the
        // domain reads cell 0 and a level out of range the nearest one.  The
        // values are used only where the lane and level are in range (the
        // guards of the phases below), and no defaults are kept per lane.
        auto load_coef = [&]() {
            const unsigned l2 = (unsigned)max(L2, 0), l1 = (unsigned)max(0, min(L1, nz - 2));
#pragma unroll
            for (int t = 0; t < OWN; t++) {
                {
                    const unsigned c = own_base[t] + sk * l2;
                    gcfloat *v4 = gx + 4 * (size_t)c;
                    const float4 v = make_float4(v4[0], v4[1], v4[2], v4[3]);
                    kx[t][0] = v.x;
                    kx[t][1] = v.y;
                    kx[t][2] = v.z;
                    kx[t][3] = v.w;
                    kg[t][0] = Gp[c];
                    kg[t][1] = Gzp[c];
                }
                {
                    const unsigned c = own_base[t] + sk * l1;
                    gcfloat *v4 = gz + 4 * (size_t)c;
                    const float4 v = make_float4(v4[0], v4[1], v4[2], v4[3]);
                    kz[t][0] = v.x;
                    kz[t][1] = v.y;
                    kz[t][2] = v.z;
                    kz[t][3] = v.w;
                    kz[t][4] = dzp[c];
                }
            }
#pragma unroll
            for (int r = 0; r < SR; r++) {
                const unsigned c = rowbase[r] + sk * l2;
                gcfloat *v4 = gy + 4 * (size_t)c;
                const float4 v = make_float4(v4[0], v4[1], v4[2], v4[3]);
                ky[r][0] = v.x;
                ky[r][1] = v.y;
                ky[r][2] = v.z;
                ky[r][3] = v.w;
            }
        };
        if (PF) load_coef();
        __syncthreads();
        if (!PF) load_coef();

        // Phases B, C and D divide as the reference does, one IEEE quotient
        // per division.  With FD each quotient is the three-operation one of
        // fdiv.hpp, which is the same value wherever its operands are in
        // range; a guard collects the operands of a wave's quotients, and a
        // wave with an operand out of range computes the phase again with
        // the IEEE division before anything is written.
        using fast_t = std::integral_constant<bool, true>;
        using ieee_t = std::integral_constant<bool, false>;

        // ---------------- (B) unlimited pseudo-velocities (mpdata_fluxes)
        // (the x and y faces go to LDS as they are computed: a phase computed
        // again writes the same places; the z faces are kept until the phase
        // is accepted, since they shift the column's pipeline)
        float zw[OWN], zf[OWN];
        auto phase_b = [&](auto fd_c) -> bool {
            constexpr bool F = decltype(fd_c)::value;
            DivGuard g;
            auto dv = [&](float x, float d) { return fdiv<F>(x, d, g); };
            // z face on top of level L1 for the owned columns (:211-246, :380)
            if (L1 >= 0 && L1 <= nz - 1) {
                const bool kz_ok = L1 <= nz - 2;
#pragma unroll
                for (int t = 0; t < OWN; t++) {
                    const int row = ly + NW * (rb + t);
                    const int gj = j0 + row;
                    float wv = 0.0f, fv = 0.0f;
                    if (kz_ok && lane_out && own_in[t]) {
                        const float(*qa)[CX] = sq1[L1 & 3];  // level k
                        const float(*qb)[CX] = sq1[m & 3];   // level k+1
                        const float l = qa[row][lx], r = qb[row][lx];
                        float val = dv(kz[t][0] * (r - l), r + l + eps);
                        const bool cu = gi >= 1 && gi <= nx - 2, cv = gj >= 1 && gj <= ny - 2;
                        const float ys = (F && (cu || cv)) ? recip_of(kz[t][3], g) : 0.0f;
                        if (cu) {  // WxU (:227-233)
                            const float eq = dv(qb[row][lx + 1] - qa[row][lx - 1] + qa[row][lx + 1] - qb[row][lx - 1],
                                                qa[row][lx + 1] + qb[row][lx + 1] + qa[row][lx - 1] + qb[row][lx - 1] + eps);
                            val = val - fdiv_y<F>(kz[t][1] * eq, kz[t][3], ys, g);
                        }
                        if (cv) {  // WxV (:238-245)
                            const float eq = dv(qb[row + 1][lx] - qa[row - 1][lx] + qa[row + 1][lx] - qb[row - 1][lx],
                                                qa[row + 1][lx] + qb[row - 1][lx] + qb[row + 1][lx] + qa[row - 1][lx] + eps);
                            val = val - fdiv_y<F>(kz[t][2] * eq, kz[t][3], ys, g);
                        }
                        wv = val * 0.5f * kz[t][4];
                        fv = flux1(l, r, wv);
                    }
                    zw[t] = wv;
                    zf[t] = fv;
                }
            }
            if (L2 >= 0) {
                const float(*q1m)[CX] = sq1[L2 & 3];
                const float(*q1u)[CX] = sq1[(L2 + 1) & 3];
                const float(*q1d)[CX] = sq1[(L2 + 3) & 3];
                const bool kin = L2 >= 1 && L2 <= nz - 2;
                // x faces (left of column gi) on the owned rows, lanes 1..63 (:131-165, :378)
#pragma unroll
                for (int t = 0; t < OWN; t++) {
                    const int row = ly + NW * (rb + t);
                    const int gj = j0 + row;
                    float uv = 0.0f, fv = 0.0f;
                    if (lx >= 1 && gi >= 1 && own_in[t]) {
                        const float r = q1m[row][lx], l = q1m[row][lx - 1];
                        float val = dv(kx[t][0] * (r - l), r + l + eps);
                        const bool cv = gj >= 1 && gj <= ny - 2;
                        const float ys = (F && (cv || kin)) ? recip_of(kx[t][3], g) : 0.0f;
                        if (cv) {  // UxV (:145-152)
                            const float eq = dv(q1m[row + 1][lx] - q1m[row - 1][lx] + q1m[row + 1][lx - 1] - q1m[row - 1][lx - 1],
                                                q1m[row + 1][lx] + q1m[row - 1][lx] + q1m[row + 1][lx - 1] + q1m[row - 1][lx - 1] + eps);
                            val = val - fdiv_y<F>(kx[t][1] * eq, kx[t][3], ys, g);
                        }
                        if (kin) {  // UxW (:157-164)
                            const float eq = dv(q1u[row][lx] - q1d[row][lx] + q1u[row][lx - 1] - q1d[row][lx - 1],
                                                q1u[row][lx] + q1d[row][lx] + q1u[row][lx - 1] + q1d[row][lx - 1] + eps);
                            val = val - fdiv_y<F>(kx[t][2] * eq, kx[t][3], ys, g);
                        }
                        uv = val * 0.5f;
                        fv = flux1(l, r, uv);
                    }
                    sxu[row][lx] = uv;
                    sxf[row][lx] = fv;
                }
                // y faces (south of row gj) on the staged rows 1..19, lanes 2..61 (:169-205, :379)
#pragma unroll
                for (int r4 = 0; r4 < SR; r4++) {
                    const int row = ly + NW * r4;
                    if (r4 >= nsr || row < 1) continue;
                    const int gj = j0 + row;
                    float vv = 0.0f, fv = 0.0f;
                    if (lane_out && gj >= 1 && gj <= ny - 1) {
                        const float r = q1m[row][lx], l = q1m[row - 1][lx];
                        float val = dv(ky[r4][0] * (r - l), r + l + eps);
                        const bool cu = gi >= 1 && gi <= nx - 2;
                        const float ys = (F && (cu || kin)) ? recip_of(ky[r4][3], g) : 0.0f;
                        if (cu) {  // VxU (:186-192)
                            const float eq = dv(q1m[row - 1][lx + 1] - q1m[row][lx - 1] + q1m[row][lx + 1] - q1m[row - 1][lx - 1],
                                                q1m[row][lx + 1] + q1m[row - 1][lx + 1] + q1m[row][lx - 1] + q1m[row - 1][lx - 1] + eps);
                            val = val - fdiv_y<F>(ky[r4][1] * eq, ky[r4][3], ys, g);
                        }
                        if (kin) {  // VxW (:197-204)
                            const float eq = dv(q1u[row - 1][lx] - q1d[row][lx] + q1u[row][lx] - q1d[row - 1][lx],
                                                q1u[row - 1][lx] + q1d[row][lx] + q1u[row][lx] + q1d[row - 1][lx] + eps);
                            val = val - fdiv_y<F>(ky[r4][2] * eq, ky[r4][3], ys, g);
                        }
                        vv = val * 0.5f;
                        fv = flux1(l, r, vv);
                    }
                    syu[row][lx] = vv;
                    syf[row][lx] = fv;
                }
            }
            return guard_ok(g);
        };
        if (!FD || __any(!phase_b(fast_t{}))) phase_b(ieee_t{});
        if (L1 >= 0 && L1 <= nz - 1) {
#pragma unroll
            for (int t = 0; t < OWN; t++) {
                wu2[t] = wu1[t];
                fz2[t] = fz1[t];
                wu1[t] = zw[t];
                fz1[t] = zf[t];
            }
        }
        __syncthreads();

        // ---------------- (C) FCT: per-cell betas (adv_mpdata_FCT_core.f90:47-116)
        float Fzc[OWN];  // flux on top of level L2
#pragma unroll
        for (int t = 0; t < OWN; t++) Fzc[t] = 0.0f;
        if (FCT) {
            float zbi[OWN], zbo[OWN], fzc[OWN];  // the x and y betas go to LDS directly
            auto phase_c = [&](auto fd_c) -> bool {
                constexpr bool F = decltype(fd_c)::value;
                DivGuard g;
                auto dv = [&](float x, float d) { return fdiv<F>(x, d, g); };
                if (L2 >= 0) {
                    const float(*q1m)[CX] = sq1[L2 & 3];
                    const float(*q0m)[CX] = sq0[L2 % 3];
                    // x cells of the owned rows, lanes 1..62
#pragma unroll
                    for (int t = 0; t < OWN; t++) {
                        const int row = ly + NW * (rb + t);
                        if (lx >= 1 && lx <= CX - 2 && lane_in) {
                            float qmax, qmin, fin = 0.0f, fout = 0.0f;
                            const float qc1 = q1m[row][lx];
                            if (gi == 0) {
                                qmax = fmaxf(fmaxf(qc1, q1m[row][lx + 1]), fmaxf(q0m[row][lx], q0m[row][lx + 1]));
                                qmin = fminf(fminf(qc1, q1m[row][lx + 1]), fminf(q0m[row][lx], q0m[row][lx + 1]));
                            } else if (gi == nx - 1) {
                                qmax = max3(q1m[row][lx - 1], qc1, q0m[row][lx - 1]);
                                qmin = min3(q1m[row][lx - 1], qc1, q0m[row][lx - 1]);
                            } else {
                                qmax = fmaxf(max3(q1m[row][lx - 1], qc1, q1m[row][lx + 1]),
                                             max3(q0m[row][lx - 1], q0m[row][lx], q0m[row][lx + 1]));
                                qmin = fminf(min3(q1m[row][lx - 1], qc1, q1m[row][lx + 1]),
                                             min3(q0m[row][lx - 1], q0m[row][lx], q0m[row][lx + 1]));
                                const float fl = sxf[row][lx], fr = sxf[row][lx + 1];
                                fin = fmaxf(0.0f, fl) - fminf(0.0f, fr);
                                fout = fmaxf(0.0f, fr) - fminf(0.0f, fl);
                            }
                            sbxi[row][lx] = dv(qmax - qc1, fin + eps15);
                            sbxo[row][lx] = dv(qc1 - qmin, fout + eps15);
                        }
                    }
                    // y cells on the staged rows 1..18, lanes 2..61
#pragma unroll
                    for (int r4 = 0; r4 < SR; r4++) {
                        const int row = ly + NW * r4;
                        if (r4 >= nsr || row < 1 || row > CY - 2) continue;
                        const int gj = j0 + row;
                        if (lane_out && gj >= 0 && gj < ny) {
                            float qmax, qmin, fin = 0.0f, fout = 0.0f;
                            const float qc1 = q1m[row][lx];
                            if (gj == 0) {
                                qmax = fmaxf(fmaxf(qc1, q1m[row + 1][lx]), fmaxf(q0m[row][lx], q0m[row + 1][lx]));
                                qmin = fminf(fminf(qc1, q1m[row + 1][lx]), fminf(q0m[row][lx], q0m[row + 1][lx]));
                            } else if (gj == ny - 1) {
                                qmax = max3(q1m[row - 1][lx], qc1, q0m[row - 1][lx]);
                                qmin = min3(q1m[row - 1][lx], qc1, q0m[row - 1][lx]);
                            } else {
                                qmax = fmaxf(max3(q1m[row - 1][lx], qc1, q1m[row + 1][lx]),
                                             max3(q0m[row - 1][lx], q0m[row][lx], q0m[row + 1][lx]));
                                qmin = fminf(min3(q1m[row - 1][lx], qc1, q1m[row + 1][lx]),
                                             min3(q0m[row - 1][lx], q0m[row][lx], q0m[row + 1][lx]));
                                const float fs = syf[row][lx], fn = syf[row + 1][lx];
                                fin = fmaxf(0.0f, fs) - fminf(0.0f, fn);
                                fout = fmaxf(0.0f, fn) - fminf(0.0f, fs);
                            }
                            sbyi[row][lx] = dv(qmax - qc1, fin + eps15);
                            sbyo[row][lx] = dv(qc1 - qmin, fout + eps15);
                        }
                    }
                }
                // z: betas of cell L1 of each owned column, then the limited
                // velocity and flux on the face between L2 and L1 (flux_is_w ends)
                if (L1 >= 0 && L1 <= nz - 1) {
#pragma unroll
                    for (int t = 0; t < OWN; t++) {
                        const int row = ly + NW * (rb + t);
                        const float qc1 = sq1[L1 & 3][row][lx];
                        float qmax, qmin, fin, fout;
                        if (L1 == 0) {
                            const float qu1 = sq1[m & 3][row][lx];
                            qmax = fmaxf(fmaxf(qc1, qu1), fmaxf(z0b[t], z0c[t]));
                            qmin = fminf(fminf(qc1, qu1), fminf(z0b[t], z0c[t]));
                            fin = 0.0f - fminf(0.0f, fz1[t]);
                            fout = fmaxf(0.0f, fz1[t]);
                        } else if (L1 == nz - 1) {
                            const float qd1 = sq1[L2 & 3][row][lx];
                            qmax = max3(qd1, qc1, z0a[t]);
                            qmin = min3(qd1, qc1, z0a[t]);
                            fin = fmaxf(0.0f, fz2[t]) - fminf(0.0f, fz2[t]);
                            fout = fmaxf(0.0f, fz2[t]) - fminf(0.0f, fz2[t]);
                        } else {
                            const float qd1 = sq1[L2 & 3][row][lx], qu1 = sq1[m & 3][row][lx];
                            qmax = fmaxf(max3(qd1, qc1, qu1), max3(z0a[t], z0b[t], z0c[t]));
                            qmin = fminf(min3(qd1, qc1, qu1), min3(z0a[t], z0b[t], z0c[t]));
                            fin = fmaxf(0.0f, fz2[t]) - fminf(0.0f, fz1[t]);
                            fout = fmaxf(0.0f, fz1[t]) - fminf(0.0f, fz2[t]);
                        }
                        const float bi = dv(qmax - qc1, fin + eps15);
                        const float bo = dv(qc1 - qmin, fout + eps15);
                        fzc[t] = 0.0f;
                        if (L2 >= 0 && L2 <= nz - 2) {
                            float wl = wu2[t];
                            if (wl > 0.0f)
                                wl = fminf(fminf(1.0f, bi), bzo[t]) * wl;
                            else if (wl < 0.0f)
                                wl = fminf(fminf(1.0f, bzi[t]), bo) * wl;
                            fzc[t] = flux1(sq1[L2 & 3][row][lx], qc1, wl);
                        }
                        zbi[t] = bi;
                        zbo[t] = bo;
                    }
                }
                return guard_ok(g);
            };
            if (!FD || __any(!phase_c(fast_t{}))) phase_c(ieee_t{});
            if (L1 >= 0 && L1 <= nz - 1) {
#pragma unroll
                for (int t = 0; t < OWN; t++) {
                    Fzc[t] = fzc[t];
                    bzi[t] = zbi[t];
                    bzo[t] = zbo[t];
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < OWN; t++) Fzc[t] = (L2 >= 0 && L2 <= nz - 2) ? fz2[t] : 0.0f;
        }
        __syncthreads();

        // ---------------- (D) flux divergence on level L2 (upwind_advection
        // :86-97) with the limited velocities of the cell's four side faces;
        // a face shared by two cells is limited identically by both
        if (L2 >= 0) {
            const float(*q1m)[CX] = sq1[L2 & 3];
            float qn[OWN];
            auto phase_d = [&](auto fd_c) -> bool {
                constexpr bool F = decltype(fd_c)::value;
                DivGuard g;
                auto dv = [&](float x, float d) { return fdiv<F>(x, d, g); };
#pragma unroll
                for (int t = 0; t < OWN; t++) {
                    const int row = ly + NW * (rb + t);
                    const int gj = j0 + row;
                    qn[t] = 0.0f;
                    if (lane_out && own_in[t] && L2 >= k0) {
                        const float qc1 = q1m[row][lx];
                        float q = qc1;
                        if (!(gi == 0 || gi == nx - 1 || gj == 0 || gj == ny - 1)) {
                            float fl, fr, fs, fn;
                            if (FCT) {
                                float u = sxu[row][lx];  // face between lx-1 and lx
                                if (u > 0.0f)
                                    u = fminf(fminf(1.0f, sbxi[row][lx]), sbxo[row][lx - 1]) * u;
                                else if (u < 0.0f)
                                    u = fminf(fminf(1.0f, sbxi[row][lx - 1]), sbxo[row][lx]) * u;
                                fl = flux1(q1m[row][lx - 1], qc1, u);
                                u = sxu[row][lx + 1];  // face between lx and lx+1
                                if (u > 0.0f)
                                    u = fminf(fminf(1.0f, sbxi[row][lx + 1]), sbxo[row][lx]) * u;
                                else if (u < 0.0f)
                                    u = fminf(fminf(1.0f, sbxi[row][lx]), sbxo[row][lx + 1]) * u;
                                fr = flux1(qc1, q1m[row][lx + 1], u);
                                float v = syu[row][lx];  // face between row-1 and row
                                if (v > 0.0f)
                                    v = fminf(fminf(1.0f, sbyi[row][lx]), sbyo[row - 1][lx]) * v;
                                else if (v < 0.0f)
                                    v = fminf(fminf(1.0f, sbyi[row - 1][lx]), sbyo[row][lx]) * v;
                                fs = flux1(q1m[row - 1][lx], qc1, v);
                                v = syu[row + 1][lx];  // face between row and row+1
                                if (v > 0.0f)
                                    v = fminf(fminf(1.0f, sbyi[row + 1][lx]), sbyo[row][lx]) * v;
                                else if (v < 0.0f)
                                    v = fminf(fminf(1.0f, sbyi[row][lx]), sbyo[row + 1][lx]) * v;
                                fn = flux1(qc1, q1m[row + 1][lx], v);
                            } else {
                                fl = sxf[row][lx];
                                fr = sxf[row][lx + 1];
                                fs = syf[row][lx];
                                fn = syf[row + 1][lx];
                            }
                            if (D.compensated)
                                q = divergence_compensated(qc1, fl, fr, fs, fn, Fzc[t], Fzp[t],
                                                           kg[t][0], kg[t][1]);
                            else {
                                q = q - dv((fr - fl) + (fn - fs), kg[t][0]);
                                q = q - dv(Fzc[t] - Fzp[t], kg[t][1]);
                            }
                        }
                        qn[t] = q;
                    }
                }
                return guard_ok(g);
            };
            if (!FD || __any(!phase_d(fast_t{}))) phase_d(ieee_t{});
#pragma unroll
            for (int t = 0; t < OWN; t++) {
                if (lane_out && own_in[t] && L2 >= k0) {
                    const unsigned c = own_base[t] + sk * (unsigned)L2;
                    st_w<QK>(p.out, c, p.kindo, p.inv_scaleo, qn[t],
                             is_sr(p.kindo) ? sr_bits(c, s, tstep) : 0u);
                }
                Fzp[t] = Fzc[t];
            }
        }
    }
}

// The same ring conditions for every species of every domain of a batch: one
// launch for the x sides and one for the y sides (the y sides copy the
// corners the x sides have written), each over a table of (domain, species)
// entries with the species' current buffer; a block row per entry.  The
// conditions copy the stored bits, so one kernel serves every storage kind.
struct BcEntry {
    Dims d;
    void *q;
    int esz;  // bytes per element: 4 (FP32) or 2 (16-bit kinds)
};

template <typename E>
__device__ __forceinline__ void bc_x_cells(E *q, const Dims &d, int k, int j, int bc_x)
{
    if (bc_x == 1) {
        q[ix(d, d.nx - 1, k, j)] = q[ix(d, d.nx - 2, k, j)];
    } else {
        q[ix(d, 0, k, j)] = q[ix(d, d.nx - 2, k, j)];
        q[ix(d, d.nx - 1, k, j)] = q[ix(d, 1, k, j)];
    }
}

template <typename E>
__device__ __forceinline__ void bc_y_cells(E *q, const Dims &d, int i, int k)
{
    q[ix(d, i, k, 0)] = q[ix(d, i, k, d.ny - 2)];
    q[ix(d, i, k, d.ny - 1)] = q[ix(d, i, k, 1)];
}

__global__ void __launch_bounds__(256) k_bc_x_all(const BcEntry *__restrict__ tab, int bc_x)
{
    const BcEntry e = tab[blockIdx.y];
    const unsigned t = blockIdx.x * 256u + threadIdx.x;
    if (t >= (unsigned)e.d.ny * (unsigned)e.d.nz) return;
    const int k = (int)(t % (unsigned)e.d.nz), j = (int)(t / (unsigned)e.d.nz);
    if (e.esz == 4)
        bc_x_cells(static_cast<uint32_t *>(e.q), e.d, k, j, bc_x);
    else
        bc_x_cells(static_cast<uint16_t *>(e.q), e.d, k, j, bc_x);
}

__global__ void __launch_bounds__(256) k_bc_y_all(const BcEntry *__restrict__ tab)
{
    const BcEntry e = tab[blockIdx.y];
    const unsigned t = blockIdx.x * 256u + threadIdx.x;
    if (t >= (unsigned)e.d.nx * (unsigned)e.d.nz) return;
    const int i = (int)(t % (unsigned)e.d.nx), k = (int)(t / (unsigned)e.d.nx);
    if (e.esz == 4)
        bc_y_cells(static_cast<uint32_t *>(e.q), e.d, i, k);
    else
        bc_y_cells(static_cast<uint16_t *>(e.q), e.d, i, k);
}

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct FastState {
    Dims d{};
    size_t N = 0;
    std::string err;
    std::vector<FastSpecies> sp;         // home buffers as bound
    std::vector<void *> alt;             // second home buffer per species
    int cur = 0;                         // which home holds the current state
    int gen = 0;                         // bumped whenever the species buffers change
    float *s1 = nullptr, *s2 = nullptr;  // FP32 scratch, nspec * N each
    float *coef = nullptr;               // NCOEF values per cell (float4 groups)
    SpecPass *d_tab = nullptr;           // [2 parities][passes][nspec]
    int tab_order = -1;
    size_t tab_cap = 0;
    unsigned *d_step = nullptr;          // steps taken (stochastic rounding counter)
};

struct FastBatch {
    std::vector<FastDomain> doms;
    std::vector<int> gens;
    int order = 0, fct = 0;
    DomDesc *d_desc = nullptr;           // [2 parities][n]
    std::vector<DomDesc> h_desc[2];
    std::vector<float> dts;              // dt values in the device tables
    unsigned total_cb = 0, total_work = 0;
    double cells = 0, cell_species = 0;  // for the algorithmic byte counts
    int corr_cfg = 1;    // corrective kernel: bit 0 -> 4 waves/SIMD, bit 1 -> no prefetch,
                         // bit 2 -> 60x32 tiles
    int tj = 16;         // output rows per corrective tile (tile_rows(corr_cfg))
    int nks = 1;         // requested column chunks along k
    bool tuned = false;  // launch configuration chosen (autotuned or fixed by the environment)
    bool dirty = true;   // device tables out of date
    bool has_sr = false; // a species is stored with stochastic rounding
    BcEntry *d_bc = nullptr;  // ring-condition table: [2 parities][n_bc] (fast_batch_bc)
    int n_bc = 0;
    unsigned bc_gx = 0, bc_gy = 0;  // blocks per entry for the x and y sides
    bool q16 = false;    // a species is stored in 16 bits
    bool f16 = false;    // a species is stored in FP16 (kernels for any kind)
};

namespace {

size_t kind_bytes(int kind) { return kind == ICAR_ADV_STORE_F32 ? 4 : 2; }

void free_species_buffers(FastState *st)
{
    for (void *p : st->alt) (void)hipFree(p);
    st->alt.clear();
    (void)hipFree(st->s1);
    (void)hipFree(st->s2);
    st->s1 = st->s2 = nullptr;
}

void *home(FastState *st, int s, int which)
{
    return which == 0 ? st->sp[s].dev : st->alt[s];
}

// Per-pass species tables of one domain for both home parities.
int build_tables(FastState *st, int order)
{
    const int S = (int)st->sp.size();
    const int P = order < 1 ? 1 : order;
    const size_t need = (size_t)2 * P * S;
    // FP32 buffers of the intermediate passes: s1 from order 2 on, s2 from
    // order 3 on (at order 1 the donor-cell pass writes the result directly)
    if (P >= 2 && !st->s1 && hipMalloc(&st->s1, (size_t)S * st->N * sizeof(float)) != hipSuccess) {
        st->s1 = nullptr;
        st->err = "hipMalloc of the FP32 pass buffer failed";
        return ICAR_ADV_ENOMEM;
    }
    if (P >= 3 && !st->s2 && hipMalloc(&st->s2, (size_t)S * st->N * sizeof(float)) != hipSuccess) {
        st->s2 = nullptr;
        st->err = "hipMalloc of the second FP32 pass buffer failed";
        return ICAR_ADV_ENOMEM;
    }
    if (need > st->tab_cap) {
        (void)hipFree(st->d_tab);
        st->d_tab = nullptr;
        if (hipMalloc(&st->d_tab, need * sizeof(SpecPass)) != hipSuccess) {
            st->err = "hipMalloc of the species pass tables failed";
            st->tab_cap = 0;
            return ICAR_ADV_ENOMEM;
        }
        st->tab_cap = need;
    }
    std::vector<SpecPass> h(need);
    for (int par = 0; par < 2; par++)
        for (int pass = 1; pass <= P; pass++)
            for (int s = 0; s < S; s++) {
                const FastSpecies &f = st->sp[s];
                const float sc = ldexpf(1.0f, f.pow2), isc = ldexpf(1.0f, -f.pow2);
                float *sa = st->s1 ? st->s1 + (size_t)s * st->N : nullptr;
                float *sb = st->s2 ? st->s2 + (size_t)s * st->N : nullptr;
                SpecPass e{};
                void *hcur = home(st, s, par), *hoth = home(st, s, 1 - par);
                if (pass == 1) {  // q0 = home
                    e.q0 = hcur, e.kind0 = f.storage, e.scale0 = sc, e.q1 = nullptr;
                } else if (pass == 2) {  // q0 = home, q1 = first pass
                    e.q0 = hcur, e.kind0 = f.storage, e.scale0 = sc, e.q1 = sa;
                } else {  // q0 = q1 = previous pass (adv_mpdata.f90:388-397)
                    e.q0 = (pass % 2 == 1) ? sb : sa;
                    e.kind0 = ICAR_ADV_STORE_F32, e.scale0 = 1.0f, e.q1 = nullptr;
                }
                if (pass == P) {  // home is not read after pass 2
                    e.out = (P <= 2) ? hoth : hcur;
                    e.kindo = f.storage, e.inv_scaleo = isc;
                } else {
                    e.out = (pass % 2 == 1) ? sa : sb;
                    e.kindo = ICAR_ADV_STORE_F32, e.inv_scaleo = 1.0f;
                }
                h[((size_t)par * P + (pass - 1)) * S + s] = e;
            }
    if (hipMemcpy(st->d_tab, h.data(), need * sizeof(SpecPass), hipMemcpyHostToDevice) !=
        hipSuccess) {
        st->err = "upload of the species pass tables failed";
        return ICAR_ADV_ENOMEM;
    }
    st->tab_order = order;
    return 0;
}

unsigned cell_blocks(const Dims &d)
{
    return (unsigned)((d.nx + 63) / 64) * (unsigned)d.nz * (unsigned)((d.ny + 3) / 4);
}

}  // namespace

FastState *fast_create(Dims d)
{
    FastState *st = new FastState();
    st->d = d;
    st->N = (size_t)d.nx * d.ny * d.nz;
    if (hipMalloc(&st->coef, (size_t)NCOEF * st->N * sizeof(float)) != hipSuccess ||
        hipMalloc(&st->d_step, sizeof(unsigned)) != hipSuccess ||
        hipMemset(st->d_step, 0, sizeof(unsigned)) != hipSuccess) {
        fast_destroy(st);
        return nullptr;
    }
    return st;
}

void fast_destroy(FastState *st)
{
    if (!st) return;
    free_species_buffers(st);
    (void)hipFree(st->coef);
    (void)hipFree(st->d_tab);
    (void)hipFree(st->d_step);
    delete st;
}

int fast_set_species(FastState *st, const FastSpecies *sp, int n, hipStream_t stream)
{
    // At mpdata_order <= 2 the steps alternate between the two home buffers.
    // When the current state sits in the second buffer, a species that keeps
    // its home buffer (same buffer, same storage) gets its current values
    // copied home first, so that binding another species leaves it as it was.
    if (st->cur == 1) {
        for (int s = 0; s < n; s++)
            for (size_t o = 0; o < st->sp.size() && o < st->alt.size(); o++)
                if (st->sp[o].dev == sp[s].dev && st->sp[o].storage == sp[s].storage &&
                    st->alt[o]) {
                    if (hipMemcpyAsync(sp[s].dev, st->alt[o],
                                       st->N * kind_bytes(sp[s].storage), hipMemcpyDeviceToDevice,
                                       stream) != hipSuccess) {
                        st->err = "copy of a species' current state to its home buffer failed";
                        return ICAR_ADV_ENOMEM;
                    }
                    break;
                }
        if (hipStreamSynchronize(stream) != hipSuccess) {
            st->err = "copy of the species' current states to their home buffers failed";
            return ICAR_ADV_ENOMEM;
        }
    }
    free_species_buffers(st);
    st->sp.assign(sp, sp + n);
    st->cur = 0;
    st->tab_order = -1;
    st->gen++;
    if (n == 0) return 0;
    st->alt.assign(n, nullptr);
    for (int s = 0; s < n; s++)
        if (hipMalloc(&st->alt[s], st->N * kind_bytes(sp[s].storage) + 64) != hipSuccess) {
            st->err = "hipMalloc of a species buffer failed";
            return ICAR_ADV_ENOMEM;
        }
    return 0;  // the FP32 pass buffers follow the order (build_tables)
}

void *fast_species_ptr(FastState *st, int s)
{
    if (!st || s < 0 || s >= (int)st->sp.size()) return nullptr;
    return home(st, s, st->cur);
}

int fast_parity(FastState *st) { return st ? st->cur : 0; }

void fast_set_parity(FastState *st, int parity)
{
    if (st) st->cur = parity & 1;
}

const char *fast_error(FastState *st) { return st ? st->err.c_str() : "no fused state"; }

namespace {

constexpr int MIN_CHUNK = 4;  // levels per column chunk at least

// Column chunks of every domain for about nks chunks per column, and the
// work-item layout (tile, chunk, species) that follows from them.
void set_ksplit(FastBatch *b, int nks)
{
    b->nks = nks < 1 ? 1 : nks;
    unsigned w = 0;
    for (size_t k = 0; k < b->doms.size(); k++) {
        const Dims &d = b->doms[k].st->d;
        int kch = (d.nz + b->nks - 1) / b->nks;
        if (kch < MIN_CHUNK) kch = MIN_CHUNK;
        if (kch > d.nz) kch = d.nz;
        const int chunks = (d.nz + kch - 1) / kch;
        const int tiles = ((d.nx + CTI - 1) / CTI) * ((d.ny + b->tj - 1) / b->tj);
        for (int par = 0; par < 2; par++) {
            DomDesc &E = b->h_desc[par][k];
            E.w0 = w;
            E.nks = chunks;
            E.kch = kch;
        }
        w += (unsigned)(tiles * chunks * (int)b->doms[k].st->sp.size());
    }
    b->total_work = w;
    b->dirty = true;
}

int upload_tables(FastBatch *b, std::string *err)
{
    const size_t n = b->doms.size();
    for (int par = 0; par < 2; par++)
        if (hipMemcpy(b->d_desc + (size_t)par * n, b->h_desc[par].data(), n * sizeof(DomDesc),
                      hipMemcpyHostToDevice) != hipSuccess) {
            *err = "upload of the domain table failed";
            return ICAR_ADV_ENOMEM;
        }
    b->dirty = false;
    return 0;
}

// corr_cfg: bit 0 -> 4 waves per SIMD (else 2), bit 1 -> no prefetch, bit 2 ->
// 60x32 tiles of 16 waves (which need 4 waves per SIMD: bit 0 is implied)
template <bool FCT, int QK, bool FD>
void launch_corr_t(unsigned nb, hipStream_t s, const DomDesc *D, int n, int pass, unsigned nw,
                   int cfg)
{
    constexpr unsigned T16 = Tile<16>::CNT, T32 = Tile<32>::CNT;
    switch (cfg & 7) {
    case 0: hipLaunchKernelGGL((k_corr<FCT, 2, true, QK, FD, 16>), dim3(nb), dim3(T16), 0, s, D, n, pass, nw); break;
    case 1: hipLaunchKernelGGL((k_corr<FCT, 4, true, QK, FD, 16>), dim3(nb), dim3(T16), 0, s, D, n, pass, nw); break;
    case 2: hipLaunchKernelGGL((k_corr<FCT, 2, false, QK, FD, 16>), dim3(nb), dim3(T16), 0, s, D, n, pass, nw); break;
    case 3: hipLaunchKernelGGL((k_corr<FCT, 4, false, QK, FD, 16>), dim3(nb), dim3(T16), 0, s, D, n, pass, nw); break;
    case 4:
    case 5: hipLaunchKernelGGL((k_corr<FCT, 4, true, QK, FD, 32>), dim3(nb), dim3(T32), 0, s, D, n, pass, nw); break;
    default: hipLaunchKernelGGL((k_corr<FCT, 4, false, QK, FD, 32>), dim3(nb), dim3(T32), 0, s, D, n, pass, nw); break;
    }
}

// the kernels' storage class for a batch: FP32 only, FP32 and BF16, or any kind
int storage_class(const FastBatch *b)
{
    return !b->q16 ? QK_F32 : b->f16 ? QK_ANY : QK_BF16;
}

template <bool FCT, bool FD>
void launch_corr_q(const FastBatch *b, unsigned nb, hipStream_t s, const DomDesc *D, int n,
                   int pass, unsigned nw, int cfg)
{
    switch (storage_class(b)) {
    case QK_F32: launch_corr_t<FCT, QK_F32, FD>(nb, s, D, n, pass, nw, cfg); break;
    case QK_BF16: launch_corr_t<FCT, QK_BF16, FD>(nb, s, D, n, pass, nw, cfg); break;
    default: launch_corr_t<FCT, QK_ANY, FD>(nb, s, D, n, pass, nw, cfg); break;
    }
}

int tile_rows(int cfg) { return (cfg & 4) ? 32 : 16; }

// ICAR_ADV_FAST_DIV=1 takes the guarded three-operation quotient of fdiv.hpp
// in the corrective and donor-cell kernels (the same results).  Off by
// default: its range guard costs about what the cheaper quotient saves.
bool fast_division()
{
    static const bool on = [] {
        const char *e = getenv("ICAR_ADV_FAST_DIV");
        return e && atoi(e) != 0;
    }();
    return on;
}

void launch_corr(const FastBatch *b, hipStream_t s, const DomDesc *D, int n, int pass, int cfg)
{
    const unsigned nw = b->total_work;
    const unsigned nb = (nw + 7) / 8 * 8;
    const bool fd = fast_division();
    if (b->fct) {
        if (fd) launch_corr_q<true, true>(b, nb, s, D, n, pass, nw, cfg);
        else launch_corr_q<true, false>(b, nb, s, D, n, pass, nw, cfg);
    } else {
        if (fd) launch_corr_q<false, true>(b, nb, s, D, n, pass, nw, cfg);
        else launch_corr_q<false, false>(b, nb, s, D, n, pass, nw, cfg);
    }
}

// Tuned configurations by batch shape, shared by every batch of the process.
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tuned;

std::string tune_key(const FastBatch *b)
{
    // batches of different storage classes run different kernel instantiations
    std::string k = std::to_string(b->order) + "/" + std::to_string(b->fct) + "/" +
                    std::to_string(storage_class(b));
    for (const auto &dm : b->doms)
        k += ":" + std::to_string(dm.st->d.nx) + "x" + std::to_string(dm.st->d.ny) + "x" +
             std::to_string(dm.st->d.nz) + "s" + std::to_string(dm.st->sp.size());
    return k;
}

}  // namespace

void launch_courant_b(const FastBatch *b, hipStream_t s, const DomDesc *D, int n)
{
    hipLaunchKernelGGL(k_courant_b, dim3(b->total_cb), dim3(64, 4), 0, s, D, n);
}

void launch_coef(const FastBatch *b, hipStream_t s, const DomDesc *D, int n)
{
    hipLaunchKernelGGL(k_coef, dim3(b->total_cb), dim3(64, 4), 0, s, D, n);
}

void launch_first(const FastBatch *b, hipStream_t s, const DomDesc *D, int n)
{
    const bool fd = fast_division();
    const dim3 g(b->total_cb), t(FBX, FBY);
    switch (storage_class(b)) {
    case QK_F32:
        if (fd) hipLaunchKernelGGL((k_first<QK_F32, true>), g, t, 0, s, D, n);
        else hipLaunchKernelGGL((k_first<QK_F32, false>), g, t, 0, s, D, n);
        break;
    case QK_BF16:
        if (fd) hipLaunchKernelGGL((k_first<QK_BF16, true>), g, t, 0, s, D, n);
        else hipLaunchKernelGGL((k_first<QK_BF16, false>), g, t, 0, s, D, n);
        break;
    default:
        if (fd) hipLaunchKernelGGL((k_first<QK_ANY, true>), g, t, 0, s, D, n);
        else hipLaunchKernelGGL((k_first<QK_ANY, false>), g, t, 0, s, D, n);
        break;
    }
}

int fast_batch_tune(FastBatch *b, hipStream_t s, std::string *err)
{
    if (b->tuned) return 0;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return 0;  // never time inside a graph capture: keep the defaults
    const bool fix_cfg = getenv("ICAR_ADV_CORR_CFG") != nullptr;
    const bool fix_ks = getenv("ICAR_ADV_KSPLIT") != nullptr;
    const bool cached = !fix_cfg && !fix_ks;  // results with a fixed part are not shared
    const std::string key = tune_key(b);
    if (cached) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            b->corr_cfg = it->second.first;
            b->tj = tile_rows(b->corr_cfg);
            set_ksplit(b, it->second.second);
            b->tuned = true;
            return 0;
        }
    }
    // candidates: the tile heights and launch configurations, and column
    // chunks while one chunk per column leaves the GPU with few blocks
    std::vector<int> cfgs = fix_cfg ? std::vector<int>{b->corr_cfg}
                                    : std::vector<int>{1, 0, 3, 2, 5, 7};
    const int n = (int)b->doms.size();
    const int par = b->doms[0].st->cur;
    // The timed launches are pass 2, which writes only the FP32 pass buffer
    // (order >= 3) or the alternate home buffer (order 2): both are written
    // again by the step that follows before anything reads them.  Its inputs
    // are made first, as the step makes them (Courant numbers, coefficients
    // and the donor-cell pass of the current state): timed on the real data,
    // the candidates take the path the step will take.
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 0;
    }
    int rc = 0;
    if (b->dirty && (rc = upload_tables(b, err))) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    {
        const DomDesc *D0 = b->d_desc + (size_t)par * n;
        launch_courant_b(b, s, D0, n);
        launch_coef(b, s, D0, n);
        launch_first(b, s, D0, n);
    }
    float best = 1e30f;
    int best_cfg = b->corr_cfg, best_ks = b->nks;
    for (int tj : {16, 32}) {
      std::vector<int> tcfgs;
      for (int cfg : cfgs)
          if (tile_rows(cfg) == tj) tcfgs.push_back(cfg);
      if (tcfgs.empty()) continue;
      b->tj = tj;
      std::vector<int> kss{b->nks};
      if (!fix_ks) {
          set_ksplit(b, 1);
          const unsigned base = b->total_work;
          kss = {1};
          if (base < 4096)
              for (int k : {2, 3, 4, 6, 8})
                  if ((unsigned)k * base <= 16384) kss.push_back(k);
      }
      std::vector<std::vector<int>> seen;  // chunk counts already timed
      for (int ks : kss) {
        set_ksplit(b, ks);
        std::vector<int> sig;
        for (int k = 0; k < n; k++) sig.push_back(b->h_desc[0][k].nks);
        bool dup = false;
        for (const auto &v : seen) dup = dup || v == sig;
        if (dup) continue;
        seen.push_back(sig);
        if ((rc = upload_tables(b, err))) break;
        const DomDesc *D = b->d_desc + (size_t)par * n;
        for (int cfg : tcfgs) {
            launch_corr(b, s, D, n, 2, cfg);  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 2; r++) launch_corr(b, s, D, n, 2, cfg);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) {
                *err = "autotuning launch failed";
                rc = ICAR_ADV_EINVAL;
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best * 0.97f) {  // prefer the earlier (simpler) candidate on near-ties
                best = ms;
                best_cfg = cfg;
                best_ks = ks;
            }
        }
        if (rc) break;
      }
      if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->corr_cfg = best_cfg;
    b->tj = tile_rows(best_cfg);
    set_ksplit(b, best_ks);
    b->tuned = true;
    if (cached) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = {best_cfg, best_ks};
    }
    return 0;
}

void fast_batch_config(const FastBatch *b, int *corr_cfg, int *ksplit)
{
    if (corr_cfg) *corr_cfg = b ? b->corr_cfg : -1;
    if (ksplit) *ksplit = b ? b->h_desc[0][0].nks : -1;
}

FastBatch *fast_batch_create(const FastDomain *doms, int n, int order, int fct, std::string *err)
{
    if (n < 1 || order < 1 || order > MAXP) {
        *err = "a batch needs at least one domain and 1 <= mpdata_order <= 16";
        return nullptr;
    }
    FastBatch *b = new FastBatch();
    b->doms.assign(doms, doms + n);
    b->order = order;
    b->fct = fct;
    // launch configuration: ICAR_ADV_CORR_CFG (0..7) and ICAR_ADV_KSPLIT fix
    // it, ICAR_ADV_AUTOTUNE=0 keeps the defaults; otherwise the first step
    // times the candidates on this batch
    const char *ec = getenv("ICAR_ADV_CORR_CFG"), *ek = getenv("ICAR_ADV_KSPLIT");
    const char *ea = getenv("ICAR_ADV_AUTOTUNE");
    if (ec) b->corr_cfg = atoi(ec) & 7;
    b->tj = tile_rows(b->corr_cfg);
    if (ek) b->nks = atoi(ek) < 1 ? 1 : atoi(ek);
    b->tuned = (ec && ek) || (ea && atoi(ea) == 0) || order < 2;
    unsigned cb = 0;
    for (int i = 0; i < 2; i++) b->h_desc[i].resize(n);
    for (int k = 0; k < n; k++) {
        FastState *st = doms[k].st;
        if (st->N >= 4294967296ull) {
            *err = "the fused path needs nx*ny*nz < 2^32 per domain";
            delete b;
            return nullptr;
        }
        if (st->sp.empty()) {
            *err = "a domain of the batch has no species bound";
            delete b;
            return nullptr;
        }
        if (st->tab_order != order && build_tables(st, order)) {
            *err = st->err;
            delete b;
            return nullptr;
        }
        b->gens.push_back(st->gen);
        const Dims &d = st->d;
        const int S = (int)st->sp.size();
        DomDesc D{};
        D.d = d;
        D.cb0 = cb;
        D.bx = (unsigned)((d.nx + 63) / 64);
        D.tiles_x = (d.nx + CTI - 1) / CTI;
        D.nspec = S;
        D.dx = doms[k].dx;
        D.dt = 0.0f;
        D.advect_density = doms[k].advect_density;
        D.compensated = doms[k].compensated;
        D.f = doms[k].f;
        D.sf = doms[k].sf;
        D.coef = st->coef;
        D.step = st->d_step;
        for (const FastSpecies &f : st->sp) {
            b->has_sr = b->has_sr || f.storage == ICAR_ADV_STORE_BF16_SR ||
                        f.storage == ICAR_ADV_STORE_F16_SR;
            b->q16 = b->q16 || f.storage != ICAR_ADV_STORE_F32;
            b->f16 = b->f16 || f.storage == ICAR_ADV_STORE_F16 || f.storage == ICAR_ADV_STORE_F16_SR;
        }
        cb += cell_blocks(d);
        b->cells += (double)st->N;
        b->cell_species += (double)st->N * S;
        for (int par = 0; par < 2; par++) {
            DomDesc E = D;
            for (int pass = 1; pass <= order; pass++)
                E.pass[pass - 1] = st->d_tab + ((size_t)par * order + (pass - 1)) * S;
            b->h_desc[par][k] = E;
        }
    }
    b->total_cb = cb;
    set_ksplit(b, b->nks);
    if (hipMalloc(&b->d_desc, 2 * (size_t)n * sizeof(DomDesc)) != hipSuccess) {
        *err = "hipMalloc of the domain table failed";
        fast_batch_destroy(b);
        return nullptr;
    }
    // ring-condition table: every species of every domain, at both parities
    std::vector<BcEntry> bc[2];
    for (int k = 0; k < n; k++) {
        FastState *st = doms[k].st;
        const Dims &d = st->d;
        b->bc_gx = std::max(b->bc_gx, ((unsigned)d.ny * (unsigned)d.nz + 255u) / 256u);
        b->bc_gy = std::max(b->bc_gy, ((unsigned)d.nx * (unsigned)d.nz + 255u) / 256u);
        for (int s = 0; s < (int)st->sp.size(); s++)
            for (int par = 0; par < 2; par++)
                bc[par].push_back({d, home(st, s, par), (int)kind_bytes(st->sp[s].storage)});
    }
    b->n_bc = (int)bc[0].size();
    if (b->n_bc > 0) {
        const size_t nb = (size_t)b->n_bc * sizeof(BcEntry);
        if (hipMalloc(&b->d_bc, 2 * nb) != hipSuccess ||
            hipMemcpy(b->d_bc, bc[0].data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->d_bc + b->n_bc, bc[1].data(), nb, hipMemcpyHostToDevice) != hipSuccess) {
            *err = "the ring-condition table could not be set up on the device";
            fast_batch_destroy(b);
            return nullptr;
        }
    }
    return b;
}

int fast_batch_bc(const FastBatch *b, hipStream_t s, int bc_x, int bc_y)
{
    if (b->n_bc == 0 || (bc_x == 0 && bc_y != 2)) return 0;
    const BcEntry *tab = b->d_bc + (size_t)b->doms[0].st->cur * b->n_bc;
    if (bc_x == 1 || bc_x == 2)
        hipLaunchKernelGGL(k_bc_x_all, dim3(b->bc_gx, b->n_bc), dim3(256), 0, s, tab, bc_x);
    if (bc_y == 2) hipLaunchKernelGGL(k_bc_y_all, dim3(b->bc_gy, b->n_bc), dim3(256), 0, s, tab);
    return (int)hipGetLastError();
}

void fast_batch_destroy(FastBatch *b)
{
    if (!b) return;
    (void)hipFree(b->d_desc);
    (void)hipFree(b->d_bc);
    delete b;
}

int fast_batch_prepare(FastBatch *b, const float *dt, std::string *err)
{
    const int n = (int)b->doms.size();
    for (int k = 0; k < n; k++)
        if (b->doms[k].st->gen != b->gens[k]) {
            *err = "species were re-bound after the batch was built";
            return ICAR_ADV_EINVAL;
        }
    bool same = (int)b->dts.size() == n;
    for (int k = 0; same && k < n; k++) same = b->dts[k] == dt[k];
    if (same && !b->dirty) return 0;
    b->dts.assign(dt, dt + n);
    for (int par = 0; par < 2; par++)
        for (int k = 0; k < n; k++) b->h_desc[par][k].dt = dt[k];
    return upload_tables(b, err);
}

int fast_batch_ready(FastBatch *b, hipStream_t s, const float *dt, std::string *err)
{
    int rc = fast_batch_prepare(b, dt, err);
    if (rc || b->tuned) return rc;
    if ((rc = fast_batch_tune(b, s, err))) return rc;
    return fast_batch_prepare(b, dt, err);
}

int fast_batch_advect(FastBatch *b, hipStream_t s, const float *dt, const TimedFn &timed,
                      std::string *err)
{
    int rc = fast_batch_ready(b, s, dt, err);
    if (rc) return rc;
    const int n = (int)b->doms.size();
    const int par = b->doms[0].st->cur;
    for (int k = 1; k < n; k++)
        if (b->doms[k].st->cur != par) {
            *err = "the domains of the batch are at different buffer parities";
            return ICAR_ADV_EINVAL;
        }
    const DomDesc *D = b->d_desc + (size_t)par * n;
    const double cells = b->cells, cs = b->cell_species;
    // Courant numbers: reads u, v, w, density, jacobians, dz; writes 6 fields
    timed(K_COURANT, (9.0 + 6.0) * 4 * cells, [&] { launch_courant_b(b, s, D, n); });
    // species-invariant coefficients: reads U, V, Wd, G; writes NCOEF
    timed(K_COEF, (4.0 + NCOEF) * 4 * cells, [&] { launch_coef(b, s, D, n); });
    // pass 1: donor cell for every species
    timed(K_FUSED_FIRST, 5.0 * 4 * cells + 2.0 * 4 * cs, [&] { launch_first(b, s, D, n); });
    for (int pass = 2; pass <= b->order; pass++) {
        // species read (q0 and q1 at pass 2) and written once; coefficients,
        // G, Gz and dz once
        const double bytes = (pass == 2 ? 3.0 : 2.0) * 4 * cs + (NCOEF + 3.0) * 4 * cells;
        timed(K_FUSED_CORR, bytes, [&] { launch_corr(b, s, D, n, pass, b->corr_cfg); });
    }
    if (b->has_sr)  // next step's rounding draws
        hipLaunchKernelGGL(k_tick, dim3((n + 63) / 64), dim3(64), 0, s, b->d_desc + (size_t)par * n, n);
    if (b->order <= 2)  // the results landed in the other home buffers
        for (auto &dm : b->doms) dm.st->cur = 1 - dm.st->cur;
    return 0;
}

void launch_pack(hipStream_t s, size_t n, const float *src, void *dst, int storage, int pow2);
void launch_unpack(hipStream_t s, size_t n, const void *src, float *dst, int storage, int pow2);

namespace {

__global__ void k_pack(size_t n, const float *__restrict__ src, uint16_t *__restrict__ dst,
                       int storage, float inv_scale)
{
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        dst[p] = to16(src[p], storage == ICAR_ADV_STORE_BF16_SR ? ICAR_ADV_STORE_BF16
                              : storage == ICAR_ADV_STORE_F16_SR ? ICAR_ADV_STORE_F16 : storage,
                      inv_scale, 0u);  // uploads round to nearest even
}

__global__ void k_unpack(size_t n, const uint16_t *__restrict__ src, float *__restrict__ dst,
                         int storage, float scale)
{
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        dst[p] = ld_q(src, p, storage, scale);
}

__global__ void k_bc16_x(Dims d, uint16_t *q, int bc_x)
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

__global__ void k_bc16_y(Dims d, uint16_t *q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= d.nx || k >= d.nz) return;
    q[ix(d, i, k, 0)] = q[ix(d, i, k, d.ny - 2)];
    q[ix(d, i, k, d.ny - 1)] = q[ix(d, i, k, 1)];
}


}  // namespace

void launch_pack(hipStream_t s, size_t n, const float *src, void *dst, int storage, int pow2)
{
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n)), dim3(256), 0, s, n, src,
                       static_cast<uint16_t *>(dst), storage, ldexpf(1.0f, -pow2));
}

void launch_unpack(hipStream_t s, size_t n, const void *src, float *dst, int storage, int pow2)
{
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(256), 0, s, n,
                       static_cast<const uint16_t *>(src), dst, storage, ldexpf(1.0f, pow2));
}

void launch_bc16(hipStream_t s, Dims d, void *q, int bc_x, int bc_y)
{
    uint16_t *p = static_cast<uint16_t *>(q);
    if (bc_x == 1 || bc_x == 2)
        hipLaunchKernelGGL(k_bc16_x, dim3((d.nz + 63) / 64, d.ny), dim3(64), 0, s, d, p, bc_x);
    if (bc_y == 2)
        hipLaunchKernelGGL(k_bc16_y, dim3((d.nx + 255) / 256, d.nz), dim3(256), 0, s, d, p);
}

}  // namespace icaradv
