/*
 * icar_adv.h -- C-ABI of the MI355X advection library (libicar_adv.so).
 *
 * It replaces, for the MPDATA scheme (&physics adv=2) and the standard
 * flux-form scheme with RK3 and the WRF limiter (adv=1), the three calls
 * that ICAR's advection module exposes:
 *   adv_init(domain, options)      src/physics/advection_driver.f90:26-44
 *                                  -> mpdata_init, src/physics/adv_mpdata.f90:415-442
 *                                  or adv_std_init, src/physics/advect.f90:29-71
 *   advect(domain, options, dt)    src/physics/advection_driver.f90:59-126
 *                                  (called at src/main/time_step.f90:373)
 *   adv_var_request(options)       src/physics/advection_driver.f90:46-57
 *                                  (species registry; stays on the Fortran side)
 * Plain C types only: pointers, ints, floats, doubles.  Every entry point
 * returns 0 on success, a negative ICAR_ADV_E* code for argument/shape errors,
 * or a positive hipError_t value for device errors; no C++ exception crosses
 * the boundary.  icar_adv_last_error() returns the message of the last failure
 * on a handle (or of the last failed create when the handle is NULL).
 *
 * Layout: every array is Fortran column-major with ICAR's bounds
 * (ims:ime, kms:kme, jms:jme), i fastest.  u and jacobian_u extend to ime+1,
 * v and jacobian_v to jme+1 (domain_obj.f90:1916-1931).
 *
 * Residency: species, winds, density and metric stay on the device between
 * calls.  icar_adv_upload / icar_adv_download are the only host<->device
 * synchronisation points for the data; icar_adv_advect only enqueues work on
 * the handle's stream and returns.
 *
 * Threading: one handle per process/image; calls on one handle must not be
 * concurrent (advect is called outside any OpenMP region, time_step.f90:373).
 */
#ifndef ICAR_ADV_H
#define ICAR_ADV_H

#ifdef __cplusplus
extern "C" {
#endif

/* Memory and tile bounds of the local domain, as mpdata_init copies them from
 * domain_t (adv_mpdata.f90:422-431), and the grid spacing domain%dx
 * (adv_mpdata.f90:486). */
typedef struct {
    int ims, ime, kms, kme, jms, jme;
    int its, ite, jts, jte;
    float dx;
} icar_adv_grid;

/* Namelist switches the advection path reads (options_t):
 *   mpdata_order    &adv_parameters mpdata_order   (options_obj.f90:1568,1581; default 2)
 *   mpdata_fct      &adv_parameters MPDATA_FCT     (options_obj.f90:1568,1580; default .true.)
 *   advect_density  &parameters advect_density     (options_obj.f90:1025; default .false.)
 *   debug           &parameters debug: runs the divergence check of
 *                   mpdata_compute_wind (adv_mpdata.f90:500-502) and syncs
 *   rk3             &time_parameters RK3: MPDATA under RK3 does nothing
 *                   (advection_driver.f90:112-115); advect returns 0
 *   kernel_variant  no reference counterpart: 0 = one kernel per MPDATA pass
 *                   and species, 1 = fused kernels batched over all species
 *                   of the handle (about twice as fast); both keep the reference's
 *                   IEEE FP32 operation order and give bit-identical results
 *                   for FP32 storage.  MPDATA only.
 *   advection       &physics adv (options_obj.f90:470-490): 1 the standard
 *                   flux-form scheme (adv_std, advect.f90), 2 MPDATA; 0 = 2
 *   h_order,        &adv_parameters h_order / v_order (options_obj.f90:1568,
 *   v_order         1583-1584; default 1): 1, 3 or 5 (flux3, advect.f90:
 *                   134-257).  The grid needs (h_order+1)/2 halo cells around
 *                   the tile, one more with the limiter; v_order 5 needs 4
 *                   levels.
 *   flux_corr       &adv_parameters flux_corr (:1582; default 0): 1 applies
 *                   the WRF monotonic limiter (adv_fluxcorr.f90) in the last
 *                   call of each step
 * With advection 1, rk3 selects the three-stage RK3 of advection_driver.f90:
 * 98-110 (t_factor 0.333, 0.5, 1) instead of one Euler call; the species
 * must be stored in FP32. */
typedef struct {
    int mpdata_order;
    int mpdata_fct;
    int advect_density;
    int debug;
    int rk3;
    int kernel_variant;
    int advection;
    int h_order;
    int v_order;
    int flux_corr;
} icar_adv_opts;

/* &physics adv= (icar_constants.f90:419-420); 0 in icar_adv_opts.advection
 * means MPDATA. */
enum { ICAR_ADV_STD = 1, ICAR_ADV_MPDATA = 2 };

/* Fields bound with icar_adv_bind_field (names of domain_t components). */
enum {
    ICAR_ADV_FIELD_U = 0,        /* domain%u%data_3d        (ims:ime+1, kms:kme, jms:jme) */
    ICAR_ADV_FIELD_V = 1,        /* domain%v%data_3d        (ims:ime, kms:kme, jms:jme+1) */
    ICAR_ADV_FIELD_W = 2,        /* domain%w%data_3d        mass grid */
    ICAR_ADV_FIELD_DENSITY = 3,  /* domain%density%data_3d  mass grid */
    ICAR_ADV_FIELD_JACOBIAN = 4, /* domain%jacobian         mass grid */
    ICAR_ADV_FIELD_JACOBIAN_U = 5,
    ICAR_ADV_FIELD_JACOBIAN_V = 6,
    ICAR_ADV_FIELD_JACOBIAN_W = 7,
    ICAR_ADV_FIELD_ADVECTION_DZ = 8,
    ICAR_ADV_NFIELDS = 9
};

/* Device storage of a species (the host side is always 4-byte real).  The
 * 16-bit kinds hold value * 2^-e and are rounded once per step, after the
 * FP32 passes: BF16 / F16 to nearest even; BF16_SR / F16_SR stochastically
 * (round up with probability equal to the discarded fraction of an ulp, the
 * random bits a counter-based hash of cell, species and step, so runs are
 * deterministic).  Round-to-nearest loses updates smaller than half an ulp,
 * which makes the total mass drift (about -1 % in 200 steps in BF16);
 * stochastic rounding keeps it unbiased.  Uploads round to nearest even. */
enum {
    ICAR_ADV_STORE_F32 = 0,
    ICAR_ADV_STORE_BF16 = 1,
    ICAR_ADV_STORE_F16 = 2,
    ICAR_ADV_STORE_BF16_SR = 3,
    ICAR_ADV_STORE_F16_SR = 4
};

/* Sides of the tile for the halo exchange. */
enum { ICAR_ADV_NORTH = 0, ICAR_ADV_SOUTH = 1, ICAR_ADV_EAST = 2, ICAR_ADV_WEST = 3 };

/* Error codes (negative); positive values are hipError_t. */
enum {
    ICAR_ADV_OK = 0,
    ICAR_ADV_EINVAL = -1,   /* bad argument or NULL pointer */
    ICAR_ADV_ESHAPE = -2,   /* array bounds do not match the grid */
    ICAR_ADV_EUNBOUND = -3, /* a required field or species was not bound */
    ICAR_ADV_ERANGE = -4,   /* index or size beyond a documented limit */
    ICAR_ADV_ENOMEM = -5,   /* device allocation failed */
    ICAR_ADV_ECOURANT = -6  /* debug check: divergence/Courant out of range */
};

#define ICAR_ADV_MAX_SPECIES 48
/* upload/download masks: bits 0..8 select fields, bit 16+s selects species s */
#define ICAR_ADV_MASK_FIELD(f) (1ULL << (f))
#define ICAR_ADV_MASK_ALL_FIELDS 0x1FFULL
#define ICAR_ADV_MASK_SPECIES(s) (1ULL << (16 + (s)))
#define ICAR_ADV_MASK_ALL_SPECIES 0xFFFFFFFFFFFF0000ULL

/* = adv_init/mpdata_init (advection_driver.f90:26-44, adv_mpdata.f90:415-442).
 * Validates bounds (nx, ny, nz >= 3; its, ite, jts, jte inside ims..ime x
 * jms..jme, ICAR_ADV_ESHAPE otherwise; an empty tile is allowed) and
 * allocates the device mirrors and
 * work space on the current HIP device.  Unlike mpdata_init it keeps its
 * state in the handle, so several domains can coexist in one process. */
int icar_adv_create(const icar_adv_grid *grid, const icar_adv_opts *opts, void **handle);

/* Frees every device allocation of the handle. */
int icar_adv_destroy(void *handle);

/* Registers the host array of a domain_t field (c_loc plus Fortran bounds
 * lb/ub in (i,k,j) order).  The array must outlive the handle. */
int icar_adv_bind_field(void *handle, int field_id, float *host, const int lb[3], const int ub[3]);

/* Registers advected species idx (0-based position in domain%adv_vars,
 * domain_obj.f90:68-89) with its host array (mass-grid bounds of the
 * domain).  storage selects the device format; pow2_exp is the per-species
 * power-of-two scale e: the device holds q * 2^-e (ignored for FP32). */
int icar_adv_bind_species(void *handle, int idx, float *host, int storage, int pow2_exp);

/* Copies the selected host arrays to the device (after CPU physics touched
 * them) / device arrays to the host (before microphysics, output, restart).
 * Both synchronise the handle's stream. */
int icar_adv_upload(void *handle, unsigned long long mask);
int icar_adv_download(void *handle, unsigned long long mask);

/* = advect(domain, options, dt) for adv=2: Courant numbers once
 * (mpdata_compute_wind, adv_mpdata.f90:477-504), then every bound species in
 * index order through the MPDATA passes (mpdata_advect3d, :352-413).
 * Enqueues on the handle's stream and returns. */
int icar_adv_advect(void *handle, float dt);

/* Per-species FP64 budget sum(q * jacobian * density * advection_dz) over the
 * cells advect updates (ims+1:ime-1, all k, jms+1:jme-1); per_species holds
 * one double per bound species.  Synchronises. */
int icar_adv_mass_budget(void *handle, double *per_species);

/* FP64 maxima over the interior of |U_m|, |V_m| and |W_m/dz| for the dt of
 * the last advect call (the cfl_strictness=3 measure of compute_dt,
 * time_step.f90:98-127, expressed as Courant numbers): the mass-flux Courant
 * numbers of the scheme in use, MPDATA's (adv_mpdata.f90:489-498) or the
 * standard scheme's (advect.f90:444-492).  Synchronises. */
int icar_adv_max_courant(void *handle, double out[3]);

/* = test_divergence (adv_mpdata.f90:445-475), which mpdata_compute_wind
 * runs when debug is set (:500-502), on the Courant numbers of the last
 * advect call: out[0] = the number of interior cells (ims+1:ime-1, all k,
 * jms+1:jme-1) whose discrete divergence |dU + dV + dW/dz| exceeds 1e-3 (the
 * cells the reference prints a warning for), out[1] = the largest such
 * |divergence| (an FP32 value).  With debug set advect also prints the
 * count to stderr.  MPDATA only; needs one advect call first.
 * Synchronises. */
int icar_adv_divergence(void *handle, double out[2]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): icar_adv_mass_budget's FP64 sum(q * jacobian * density *
 * advection_dz) over the interior cells (ims+1:ime-1, all k, jms+1:jme-1) of
 * one FP32 species, on the caller's device arrays (mass grid, (i, k, j), i
 * fastest) with the result in *out (device memory, one double), equal bit
 * for bit to icar_adv_mass_budget's value for the same arrays.  Its per-block
 * partial sums go to a fixed array of the library on the current device
 * (ICAR_ADV_NOALLOC_PARTIALS doubles, 8 KiB, held until the process ends)
 * instead of an allocation.  It enqueues its two kernels on `stream` (a
 * hipStream_t; NULL is the null stream) and returns: no allocation, memset
 * or synchronisation, so it can be captured in a graph.  Calls on one GPU
 * share the array, so calls that overlap (on different streams) race.  The
 * first call on a device looks up the array's device address
 * (hipGetSymbolAddress): make it before capturing one.
 * The reduction takes at most 1024 blocks for any grid (each block strides
 * over its share of the cells), so every grid fits the array.
 * Errors (icar_adv_last_error(NULL) holds the message): ICAR_ADV_EINVAL for
 * nx, ny or nz below 3 or a NULL array; ICAR_ADV_ERANGE, should the
 * reduction ever take more blocks than the array holds. */
#define ICAR_ADV_NOALLOC_PARTIALS 1024
int icar_adv_mass_budget_noalloc(int nx, int ny, int nz, const float *q, const float *jacobian,
                                 const float *density, const float *advection_dz, double *out,
                                 void *stream);

/* = balance_uvw(domain, options) (src/physics/wind.f90:88-164, update
 * =.false.), which the time step runs before advect when advect_density is
 * set (src/main/time_step.f90:324): recomputes the device copy of w from
 * the device u, v, density, jacobians and advection_dz so that the mass
 * flux is non-divergent (calc_divergence with horz_only, :201-234, then
 * calc_w, :166-199), over the columns its-1..ite+1 x jts-1..jte+1 clipped to
 * the memory bounds, w = 0 elsewhere.  Uses the handle's advect_density.
 * icar_adv_download with ICAR_ADV_MASK_FIELD(ICAR_ADV_FIELD_W) returns it. */
int icar_adv_balance_w(void *handle);

/* = compute_dt(dx, u, v, w, rho, dz_levels, ..., CFL, cfl_strictness,
 * use_density=.false.) (src/main/time_step.f90:51-168, called by update_dt
 * :213-246) on the device copies of u, v and w: the reference's FP32 result
 * in *dt.  dz_levels holds the nz nominal level thicknesses
 * (options%parameters%dz_levels); cfl is cfl_reduction_factor.  Returns
 * ICAR_ADV_ECOURANT, with *dt set, when dt < 0.1 s, where the reference
 * stops.  Synchronises. */
int icar_adv_compute_dt(void *handle, float cfl, int cfl_strictness, const float *dz_levels,
                        float *dt);

/* Lateral forcing of advected species, the step that fills the outer cells
 * in ICAR's time step before the physics and advect (apply_forcing,
 * src/objects/domain_obj.f90:3457-3530, called at src/main/time_step.f90:319),
 * on the device copies.
 *
 * icar_adv_set_relax_filter = init_relax_filters (domain_obj.f90:3285-3390):
 * the relaxation filter for the domain's boundary flags (non-zero = the edge
 * is a domain boundary) and halo width (grid%halo_size).
 * icar_adv_set_forcing uploads species idx's forcing field (forcing_hi
 * data_3d) and its tendency per second (dqdt_3d), both on the mass grid, as
 * update_delta_fields sets them (:3405-3445); FP32-stored species only.
 * icar_adv_apply_forcing advances every forced species over dt seconds:
 * forcing += dqdt*dt everywhere; where the filter is 1 the species takes the
 * forcing, where it is below 1 it relaxes by filter*dt/3600 of the
 * difference (double precision intermediates, as dt%seconds() is). */
/* Halo exchange of the advected species between neighbouring subdomains
 * (domain%halo_exchange / halo_exchange_batch before advect,
 * time_step.f90:362-366; domain_obj.f90:573-650, exchangeable_obj.f90:
 * 240-392).  As in the reference, a side sends the strip inside the tile as
 * wide as its halo (rows jte-h+1..jte for the north side) over the tile's
 * range along the side, and the neighbour writes it into its halo on the
 * opposite side (rows jms..jts-1); corners are not exchanged.
 *   icar_adv_halo_bytes   bytes of one side's strip for every bound species
 *                         (species after species, each its (i, k, j) sub-box
 *                         at its storage width, 16-byte aligned)
 *   icar_adv_halo_pack    the strip into dev_buf (device memory), on the
 *                         handle's stream: what an RCCL send of this side
 *                         would carry
 *   icar_adv_halo_unpack  dev_buf into the halo on that side
 *   icar_adv_halo_exchange  from's strip on `side` straight into to's halo
 *                         on the opposite side, both handles on one device
 *                         (subdomains that share a GPU); ordered after the
 *                         sender's queued work, and the sender's later work
 *                         after the copy
 * The strips carry the stored values: both sides need the same species, the
 * same storage format per species (FP32, BF16 or FP16) with the same 2^e
 * scale, and the same strip shape.  icar_adv_halo_exchange checks this
 * (ICAR_ADV_ESHAPE); for pack and unpack it is the caller's contract. */
int icar_adv_halo_bytes(void *handle, int side, unsigned long long *bytes);
int icar_adv_halo_pack(void *handle, int side, void *dev_buf, unsigned long long bytes);
int icar_adv_halo_unpack(void *handle, int side, const void *dev_buf, unsigned long long bytes);
int icar_adv_halo_exchange(void *from, int side, void *to);

/* = domain%enforce_limits (domain_obj.f90:3092-3150, called in the last
 * updates of a step, time_step.f90:386-389): every bound species below zero
 * becomes zero on the device (-0.0 and NaN stay, as with `where (q < 0)`). */
int icar_adv_enforce_limits(void *handle);

/* Output and restart hand-off (output_obj.f90:202-230 writes the domain
 * fields): icar_adv_snapshot copies the fields and species `mask` selects
 * (the upload/download mask bits) as they are after the work queued so far,
 * first into device staging on the handle's stream (a device copy; 16-bit
 * species decoded to FP32), then into pinned host staging on a copy stream
 * that overlaps the following steps.  icar_adv_snapshot_wait blocks until
 * that copy is done and writes the values into the bound host arrays.  A
 * second snapshot, icar_adv_download or destroy completes a pending one
 * first (so a later download's values are the ones the host arrays keep);
 * icar_adv_upload of an array a pending snapshot covers returns
 * ICAR_ADV_EINVAL, since the snapshot has yet to write it. */
int icar_adv_snapshot(void *handle, unsigned long long mask);
int icar_adv_snapshot_wait(void *handle);

int icar_adv_set_relax_filter(void *handle, int west, int east, int north, int south, int halo);
int icar_adv_set_forcing(void *handle, int idx, const float *forcing_hi, const float *dqdt);
int icar_adv_apply_forcing(void *handle, double dt);

/* Message of the last failure on this handle (handle may be NULL). */
const char *icar_adv_last_error(void *handle);

/* ---- Extensions with no reference counterpart ---------------------------- */

/* Waits for all work enqueued on the handle's stream. */
int icar_adv_sync(void *handle);

/* Lateral boundary conditions on the outer ring of every species, applied on
 * the device between advect calls.  In ICAR the ring is filled by forcing and
 * halo exchange (time_step.f90:319,362-368); this does what the reference's
 * MPDATA test does (test_mpdata.f90:96-97).  bc_x: 0 fixed, 1 zero-gradient
 * outflow at i = ime, 2 periodic; bc_y: 0 fixed, 2 periodic.  x first. */
int icar_adv_apply_bc(void *handle, int bc_x, int bc_y);

/* Runs nsteps x (advect(dt), then apply_bc(bc_x, bc_y)) on the stream. */
int icar_adv_run(void *handle, float dt, int nsteps, int bc_x, int bc_y);

/* Kernel timing with HIP events recorded on the handle's stream around each
 * launch.  enable != 0 starts a fresh accumulation.  Kernel ids and names:
 * icar_adv_kernel_name(id) for id in [0, icar_adv_kernel_count()). */
int icar_adv_profile(void *handle, int enable);
int icar_adv_kernel_count(void);
const char *icar_adv_kernel_name(int id);
/* total milliseconds, number of launches and the algorithmic HBM bytes of
 * all timed launches of kernel id since profiling was enabled.  Synchronises. */
int icar_adv_kernel_stats(void *handle, int id, double *total_ms, long long *launches,
                          double *alg_bytes);

/* Launch configuration of the fused corrective kernel chosen for this handle
 * (kernel_variant 1), -1 before the first advect.  The first fused step times
 * the candidates on the handle's own grid and species count and keeps the
 * fastest (cached for the process by shape).  The tile is a column of 60 x 16
 * or 60 x 32 cells with a halo of 2, marched up through the levels; corr_cfg
 * bit 0 selects the register budget (4 waves per SIMD, else 2), bit 1 no
 * prefetch of the next level, bit 2 the 60 x 32 tile (blocks of 16 waves,
 * always 4 per SIMD); ksplit is the number of chunks each column march is
 * split into, i.e. the tile's extent in k (more blocks for small grids).  The
 * environment variables ICAR_ADV_CORR_CFG (0..7) and ICAR_ADV_KSPLIT fix
 * them, ICAR_ADV_AUTOTUNE=0 keeps 1 and 1. */
int icar_adv_launch_config(void *handle, int *corr_cfg, int *ksplit);

/* Records the nsteps loop of icar_adv_run in a HIP graph on its first call
 * and replays it on later calls with the same dt, nsteps and boundary
 * conditions (not while profiling or with debug set).  Binding a species
 * drops the recording; the next run records anew. */
int icar_adv_set_graph(void *handle, int enable);

/* The flux divergence of the fused MPDATA kernels (kernel_variant 1) with
 * compensated sums (on != 0) or in the reference's operation order (0, the
 * default): the flux differences, their sum and the two subtractions of the
 * update q - (dFx + dFy)/G - dFz/Gz (upwind_advection, adv_mpdata.f90:86-97,
 * in the corrective passes) carry their rounding errors (TwoSum), the two
 * quotients their remainders, and the errors are added before the last
 * rounding.  The result then differs from the FP32 reference in the last
 * bits; the measured effect on the error against FP64 is stated with the
 * precision tests.  Drops a recorded run; ensembles over the handle must be
 * created again (as after icar_adv_bind_species).  ICAR_ADV_EINVAL for the
 * other kernel variants and the standard scheme. */
int icar_adv_set_compensation(void *handle, int on);

/* Checks, on the current device, the three-operation quotient the kernels
 * take in place of an IEEE division where both operands are in range, when
 * the environment sets ICAR_ADV_FAST_DIV=1 (off by default) (q0 =
 * x*y, rn = q0*d - x, q = q0 - rn*y with y the hardware reciprocal of d
 * refined by one Newton step): for every divisor significand in [d_first,
 * d_first + d_count) (d in [1, 2); d_first + d_count <= 2^23) and every
 * numerator significand (x in [1, 2)) it must equal x / d bit for bit, the
 * reciprocal must scale exactly with d's exponent over [-60, 40], and a
 * negative numerator must give the negated quotient.  *mismatches gets the
 * number of failures, *first the first found ((d bits << 32) | x bits, or
 * all ones).  All 2^23 divisors take about 30 s.  Synchronises. */
int icar_adv_verify_division(unsigned d_first, unsigned d_count, unsigned long long *mismatches,
                             unsigned long long *first);

/* ---- Ensembles ------------------------------------------------------------
 * No reference counterpart (one ICAR image holds one domain): a set of
 * independent domains of uneven sizes advanced together, one launch per
 * pass covering every (domain, tile, species), the step loop optionally
 * recorded in a HIP graph.  All handles need kernel_variant 1, RK3 off and
 * the same mpdata_order and MPDATA_FCT; the ensemble uses their device
 * state, so call icar_adv_ensemble_sync before using a member handle.
 * Members must outlive the ensemble.  Binding a member's species again (which
 * may move its device buffers) makes icar_adv_ensemble_run return
 * ICAR_ADV_EINVAL until the ensemble is created again. */
int icar_adv_ensemble_create(int n, void *const *handles, void **ens);
int icar_adv_ensemble_destroy(void *ens);
/* dt holds one time step per domain; bc_x/bc_y as icar_adv_apply_bc. */
int icar_adv_ensemble_run(void *ens, const float *dt, int nsteps, int bc_x, int bc_y,
                          int use_graph);
int icar_adv_ensemble_sync(void *ens);
const char *icar_adv_ensemble_last_error(void *ens);
int icar_adv_ensemble_profile(void *ens, int enable);
int icar_adv_ensemble_kernel_stats(void *ens, int id, double *total_ms, long long *launches,
                                   double *alg_bytes);
/* As icar_adv_launch_config, for the ensemble's batch (ksplit of member 0). */
int icar_adv_ensemble_launch_config(void *ens, int *corr_cfg, int *ksplit);

#ifdef __cplusplus
}
#endif
#endif /* ICAR_ADV_H */
