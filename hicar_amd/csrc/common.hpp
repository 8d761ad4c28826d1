// Shared definitions of the MPDATA advection library (device and host side).
//
// Index convention (identical to ICAR's Fortran column-major arrays
// (ims:ime, kms:kme, jms:jme), 0-based):  IX(i,k,j) = i + nx*(k + nz*j).
// Face fields have the mass-grid shape:
//   U[IX(i,k,j)]  x face between cells i-1 and i (slot i = 0 unused, kept 0)
//   V[IX(i,k,j)]  y face between rows j-1 and j  (slot j = 0 unused, kept 0)
//   W[IX(i,k,j)]  z face on top of level k       (k = nz-1 is the model top)
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace icaradv {

struct Dims {
    int nx, ny, nz;
};

__host__ __device__ inline size_t ix(const Dims &d, int i, int k, int j)
{
    return (size_t)i + (size_t)d.nx * ((size_t)k + (size_t)d.nz * (size_t)j);
}

// Blocks of a 1-D launch are dealt round-robin over the 8 XCDs, each with
// its own L2.  This renumbers them so that each XCD takes a contiguous range
// of the block order: blocks that read each other's rows then share an L2.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned L = blockIdx.x, B = gridDim.x;
    const unsigned q = B / 8, rem = B % 8, x = L % 8, slot = L / 8;
    return x < rem ? x * (q + 1) + slot : rem * (q + 1) + (x - rem) * q + slot;
}

// Kernel ids for the profiling accumulators (names in kernel_names()).
enum KernelId {
    K_COURANT = 0,
    K_UPWIND,
    K_PSEUDO,
    K_FCT,
    K_BC,
    K_REDUCE,
    K_COPY,
    K_FUSED_FIRST,
    K_FUSED_CORR,
    K_CONVERT,
    K_COEF,
    K_BALANCE,
    K_DT,
    K_FORCING,
    K_STD_WIND,
    K_STD_STAGE,
    K_STD_SCALE,
    K_STD_FINAL,
    K_STD_RING,
    K_HALO,
    K_CLAMP,
    K_SNAPSHOT,
    K_COUNT
};

const char *kernel_name(int id);

// Records the message of a failed call that has no handle (what
// icar_adv_last_error(NULL) returns) and returns code.
int set_call_error(int code, const char *msg);

// Per-step derived fields written by the Courant kernel (species invariant).
struct StepFields {
    float *U, *V, *W;  // Courant numbers (mpdata_compute_wind)
    float *G;          // jacobian*density            (adv_mpdata.f90:87,374)
    float *Gz;         // advection_dz*jacobian*density (adv_mpdata.f90:91-97)
    float *Wd;         // W/advection_dz              (adv_mpdata.f90:374)
};

// Raw domain inputs on the device (ICAR shapes; u/jacobian_u have nx+1 per
// row, v/jacobian_v ny+1 rows).
struct DomainFields {
    const float *u, *v, *w, *density, *jaco, *jaco_u, *jaco_v, *jaco_w, *dz;
};

// ---- exact kernels (reference IEEE operation order), kernels_exact.hip ----
void launch_courant(hipStream_t s, Dims d, float dx, float dt, int advect_density,
                    const DomainFields &f, const StepFields &o);
void launch_upwind(hipStream_t s, Dims d, const float *qin, const float *U, const float *V,
                   const float *W, const float *G, const float *Gz, float *q);
void launch_pseudo(hipStream_t s, Dims d, const float *q1, const StepFields &sf, const float *dz,
                   float *u2, float *v2, float *w2);
void launch_fct(hipStream_t s, Dims d, const float *q0, const float *q1, const float *u2,
                const float *v2, const float *w2, float *u2l, float *v2l, float *w2l);
void launch_bc(hipStream_t s, Dims d, float *q, int bc_x, int bc_y);
// FP64 mass budget partial sums: writes nblocks doubles to partial.
int mass_partials_blocks(Dims d);

// ---- wind balance and dt bound (kernels_dyn.hip) ----
void launch_balance_w(hipStream_t s, Dims d, int i_s, int i_e, int j_s, int j_e, float dx,
                      int adv_den, const DomainFields &f, float *w);
int dt_partial_blocks();
void launch_forcing(hipStream_t s, Dims d, double dt, const float *filt, float *q, float *fhi,
                    const float *dqdt);
void launch_dt_partials(hipStream_t s, Dims d, int its, int ite, int jts, int jte, float dx,
                        const float *u, const float *v, const float *w, const float *dzl,
                        float *part);
void launch_mass(hipStream_t s, Dims d, const float *q, const float *jaco, const float *density,
                 const float *dz, double *partial);
// FP64 Courant maxima partials: 3 doubles per block.  U has unx faces per
// row (nx or nx + 1); W is divided by dz unless dz is null.
void launch_courant_max(hipStream_t s, Dims d, const float *U, int unx, const float *V,
                        const float *W, const float *dz, double *partial);
void launch_copy(hipStream_t s, size_t n, const float *src, float *dst);
// Divergence check (debug): counts cells with |div| > 1e-3; out[0]=count, out[1]=max.
void launch_divergence(hipStream_t s, Dims d, const float *U, const float *V, const float *W,
                       const float *dz, double *partial);

// ---- halo strips (kernels_dyn.hip) ----
// One side's strip of every species: the box [a_i0, a_i0+ni) x [a_j0, a_j0+nj)
// of the arrays a (row length a_nx), the same-shaped box at (b_i0, b_j0) of
// the arrays b (row length b_nx), and the packed buffer (species s at byte
// offset buf_off[s], esize[s] bytes per element).
constexpr int HALO_MAX_SPECIES = 48;
struct HaloStrips {
    int n, ni, nj;
    int a_i0, a_j0, a_nx, b_i0, b_j0, b_nx;
    const void *a[HALO_MAX_SPECIES];
    void *b[HALO_MAX_SPECIES];
    void *buf;
    long long buf_off[HALO_MAX_SPECIES];
    int esize[HALO_MAX_SPECIES];
};
// mode 0: a -> buffer; 1: buffer -> b; 2: a -> b
void launch_strip(hipStream_t s, const HaloStrips &hs, int nz, int mode);
// enforce_limits on n elements of one species at its storage kind
void launch_clamp(hipStream_t s, size_t n, void *q, int storage);

// ---- standard scheme, adv=1 (kernels_std.hip) ----
// Geometry of the flux-form update: the tile (0-based memory offsets), the
// flux orders and the per-step mass fluxes and divisors of k_std_wind.
//   Um  x faces, (nx+1) per row (face i on the west side of cell i)
//   Vm  y faces, ny+1 rows (face j on the south side of row j)
//   Wm  the top face of each cell;  Rj = jacobian*rho;  Rdj = dz*jacobian*rho
struct StdGeom {
    Dims d;
    int its, ite, jts, jte;
    int horder, vorder;
    const float *Um, *Vm, *Wm, *Rj, *Rdj;
};
constexpr int STD_MAX_BATCH = 16;
// Species of one launch: the state the fluxes come from (qin), the base
// state (qold), the output, and the limiter's work arrays: dumb_q and the
// (scale_in, scale_out) pair of each cell.
struct StdBatch {
    int n;
    const float *qin[STD_MAX_BATCH];
    const float *qold[STD_MAX_BATCH];
    float *out[STD_MAX_BATCH];
    float *dq[STD_MAX_BATCH];
    float2 *sc[STD_MAX_BATCH];
};
void launch_std_wind(hipStream_t s, Dims d, float dx, float dt, int adv_den,
                     const DomainFields &f, float *Um, float *Vm, float *Wm, float *Rj,
                     float *Rdj);
// mode 0: stage; 1: stage and dumb_q; 2: dumb_q only
void launch_std_stage(hipStream_t s, const StdGeom &g, float t_factor, int mode,
                      const StdBatch &b);
void launch_std_scale(hipStream_t s, const StdGeom &g, const StdBatch &b);
void launch_std_final(hipStream_t s, const StdGeom &g, const StdBatch &b);
// the cells no stage writes: qold into out outside the tile and, with
// with_dq, into dq outside the tile plus one ring
void launch_std_ring(hipStream_t s, const StdGeom &g, const StdBatch &b, int with_dq);

}  // namespace icaradv
