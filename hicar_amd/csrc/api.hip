// C-ABI of the MPDATA advection library: handle management, host<->device
// residency, the advect() orchestration and diagnostics.
//
// advect() follows src/physics/advection_driver.f90:59-126 for adv=2:
// Courant numbers once (mpdata_compute_wind, adv_mpdata.f90:477-504), then
// each bound species through the order loop of mpdata_advect3d
// (adv_mpdata.f90:352-413).  Species are independent inside advect, so the
// fused variant advances all of them per launch with identical results.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icar_adv.h"
#include "common.hpp"
#include "fast.hpp"

using namespace icaradv;

namespace {

thread_local std::string g_create_error;

struct SpeciesSlot {
    float *host = nullptr;
    int storage = ICAR_ADV_STORE_F32;
    int pow2 = 0;
    void *dev = nullptr;  // float*, or uint16_t* for BF16/FP16 storage
    bool bound = false;
};

struct TimedLaunch {
    int id;
    hipEvent_t a, b;
    double bytes;
};

struct Handle {
    icar_adv_grid g{};
    icar_adv_opts o{};
    Dims d{};
    size_t N = 0, Nu = 0, Nv = 0;
    float *host_field[ICAR_ADV_NFIELDS] = {};
    float *dev_field[ICAR_ADV_NFIELDS] = {};
    StepFields sf{};
    float *q2 = nullptr, *u2 = nullptr, *v2 = nullptr, *w2 = nullptr;
    float *u2l = nullptr, *v2l = nullptr, *w2l = nullptr;
    float *stage = nullptr;  // FP32 staging for reduced-precision species
    SpeciesSlot sp[ICAR_ADV_MAX_SPECIES];
    int nspecies = 0;  // 1 + highest bound index
    int compact[ICAR_ADV_MAX_SPECIES] = {};  // bound-species position in the fused state
    hipStream_t stream = nullptr;
    float last_dt = 0.0f;
    bool have_dt = false;
    double *d_partial = nullptr;
    std::vector<double> h_partial;
    std::string err;
    FastState *fast = nullptr;
    // profiling
    bool prof = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double acc_ms[K_COUNT] = {};
    long long acc_n[K_COUNT] = {};
    double acc_bytes[K_COUNT] = {};
};

int set_err(Handle *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPCHK(h, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err((h), (int)e_, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

hipEvent_t pool_event(Handle *h)
{
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records an event pair around launch() when profiling is on.
template <class F>
void timed(Handle *h, int id, double bytes, F &&launch)
{
    if (!h->prof) {
        launch();
        return;
    }
    hipEvent_t a = pool_event(h), b = pool_event(h);
    if (!a || !b) {
        launch();
        return;
    }
    (void)hipEventRecord(a, h->stream);
    launch();
    (void)hipEventRecord(b, h->stream);
    h->pending.push_back({id, a, b, bytes});
}

void drain_profile(Handle *h)
{
    for (auto &t : h->pending) {
        float ms = 0.0f;
        (void)hipEventSynchronize(t.b);
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            h->acc_ms[t.id] += ms;
            h->acc_n[t.id] += 1;
            h->acc_bytes[t.id] += t.bytes;
        }
        h->event_pool.push_back(t.a);
        h->event_pool.push_back(t.b);
    }
    h->pending.clear();
}

bool grid_ok(const icar_adv_grid *g)
{
    return g->ime - g->ims + 1 >= 3 && g->jme - g->jms + 1 >= 3 && g->kme - g->kms + 1 >= 3 &&
           g->dx > 0.0f;
}

int dalloc(Handle *h, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(h, ICAR_ADV_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes,
                       hipGetErrorString(e));
    }
    return 0;
}

size_t field_elems(const Handle *h, int f)
{
    if (f == ICAR_ADV_FIELD_U || f == ICAR_ADV_FIELD_JACOBIAN_U) return h->Nu;
    if (f == ICAR_ADV_FIELD_V || f == ICAR_ADV_FIELD_JACOBIAN_V) return h->Nv;
    return h->N;
}

DomainFields domain_fields(const Handle *h)
{
    DomainFields f;
    f.u = h->dev_field[ICAR_ADV_FIELD_U];
    f.v = h->dev_field[ICAR_ADV_FIELD_V];
    f.w = h->dev_field[ICAR_ADV_FIELD_W];
    f.density = h->dev_field[ICAR_ADV_FIELD_DENSITY];
    f.jaco = h->dev_field[ICAR_ADV_FIELD_JACOBIAN];
    f.jaco_u = h->dev_field[ICAR_ADV_FIELD_JACOBIAN_U];
    f.jaco_v = h->dev_field[ICAR_ADV_FIELD_JACOBIAN_V];
    f.jaco_w = h->dev_field[ICAR_ADV_FIELD_JACOBIAN_W];
    f.dz = h->dev_field[ICAR_ADV_FIELD_ADVECTION_DZ];
    return f;
}

void free_all(Handle *h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drain_profile(h);
    for (auto e : h->event_pool) (void)hipEventDestroy(e);
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) (void)hipFree(h->dev_field[f]);
    float *bufs[] = {h->sf.U, h->sf.V, h->sf.W, h->sf.G, h->sf.Gz, h->sf.Wd, h->q2, h->u2,
                     h->v2,   h->w2,   h->u2l, h->v2l, h->w2l, h->stage};
    for (float *p : bufs) (void)hipFree(p);
    for (auto &s : h->sp) (void)hipFree(s.dev);
    (void)hipFree(h->d_partial);
    fast_destroy(h->fast);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int check_handle(void *handle, Handle **out)
{
    if (!handle) return set_err(nullptr, ICAR_ADV_EINVAL, "NULL handle");
    *out = static_cast<Handle *>(handle);
    return 0;
}

int require_bound(Handle *h)
{
    static const char *names[ICAR_ADV_NFIELDS] = {"u",          "v",          "w",
                                                  "density",    "jacobian",   "jacobian_u",
                                                  "jacobian_v", "jacobian_w", "advection_dz"};
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++)
        if (!h->dev_field[f])
            return set_err(h, ICAR_ADV_EUNBOUND, "field %s not bound", names[f]);
    return 0;
}

// Bytes a kernel must move at minimum (each array it touches once), per
// launch; used for the achieved-bandwidth figures of the profile.
double cells(const Handle *h) { return (double)h->N; }

int species_f32_ptr(Handle *h, int s, float **p)
{
    SpeciesSlot &sl = h->sp[s];
    if (sl.storage != ICAR_ADV_STORE_F32)
        return set_err(h, ICAR_ADV_EINVAL,
                       "species %d is stored in reduced precision; the exact kernels need FP32 "
                       "storage (use kernel_variant 1)",
                       s);
    *p = static_cast<float *>(sl.dev);
    return 0;
}

// Device buffer that holds the current state of species s.
void *species_dev(Handle *h, int s)
{
    if (h->fast) return fast_species_ptr(h->fast, h->compact[s]);
    return h->sp[s].dev;
}

// One species through mpdata_advect3d with the exact kernels.
int advect_species_exact(Handle *h, float *q)
{
    const Dims d = h->d;
    const int order = h->o.mpdata_order;
    const double n = cells(h);
    for (int iord = 1; iord <= order; iord++) {
        if (iord == 1) {
            timed(h, K_UPWIND, 7 * 4 * n, [&] {
                launch_upwind(h->stream, d, q, h->sf.U, h->sf.V, h->sf.W, h->sf.G, h->sf.Gz, h->q2);
            });
        } else {
            timed(h, K_PSEUDO, 9 * 4 * n, [&] {
                launch_pseudo(h->stream, d, h->q2, h->sf, h->dev_field[ICAR_ADV_FIELD_ADVECTION_DZ],
                              h->u2, h->v2, h->w2);
            });
            const float *uu = h->u2, *vv = h->v2, *ww = h->w2;
            if (h->o.mpdata_fct) {
                timed(h, K_FCT, 8 * 4 * n, [&] {
                    launch_fct(h->stream, d, q, h->q2, h->u2, h->v2, h->w2, h->u2l, h->v2l, h->w2l);
                });
                uu = h->u2l;
                vv = h->v2l;
                ww = h->w2l;
            }
            timed(h, K_UPWIND, 7 * 4 * n, [&] {
                launch_upwind(h->stream, d, h->q2, uu, vv, ww, h->sf.G, h->sf.Gz, q);
            });
        }
        if (iord != order) {
            if (iord > 1)
                timed(h, K_COPY, 2 * 4 * n, [&] { launch_copy(h->stream, h->N, q, h->q2); });
        } else if (iord == 1) {
            timed(h, K_COPY, 2 * 4 * n, [&] { launch_copy(h->stream, h->N, h->q2, q); });
        }
    }
    HIPCHK(h, hipGetLastError());
    return 0;
}

int do_advect(Handle *h, float dt)
{
    int rc = require_bound(h);
    if (rc) return rc;
    h->last_dt = dt;
    h->have_dt = true;
    const DomainFields f = domain_fields(h);
    const double n = cells(h);
    timed(h, K_COURANT, (9 + 6) * 4 * n, [&] {
        launch_courant(h->stream, h->d, h->g.dx, dt, h->o.advect_density, f, h->sf);
    });
    HIPCHK(h, hipGetLastError());
    if (h->o.debug) {
        // test_divergence (adv_mpdata.f90:500-502): the reference prints a warning
        const int nb = mass_partials_blocks(h->d);
        launch_divergence(h->stream, h->d, h->sf.U, h->sf.V, h->sf.W, f.dz, h->d_partial);
        HIPCHK(h, hipMemcpyAsync(h->h_partial.data(), h->d_partial, sizeof(double) * 2 * nb,
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        double cnt = 0, mx = 0;
        for (int b = 0; b < nb; b++) {
            cnt += h->h_partial[2 * b];
            mx = mx > h->h_partial[2 * b + 1] ? mx : h->h_partial[2 * b + 1];
        }
        if (cnt > 0)
            fprintf(stderr,
                    "icar_adv: winds are not balanced on entry to advect: %.0f cells with "
                    "|div| > 1e-3 (max %g)\n",
                    cnt, mx);
    }
    if (h->o.rk3) return 0;  // advection_driver.f90:112-115
    if (h->nspecies == 0) return 0;
    if (h->o.kernel_variant == 1) {
        rc = fast_advect(h->fast, h->stream, h->d, h->sf, f, h->o.mpdata_order, h->o.mpdata_fct,
                         [&](int id, double bytes, const std::function<void()> &fn) {
                             timed(h, id, bytes, fn);
                         });
        if (rc) return set_err(h, rc, "%s", fast_error(h->fast));
        HIPCHK(h, hipGetLastError());
        return 0;
    }
    for (int s = 0; s < h->nspecies; s++) {
        if (!h->sp[s].bound) continue;
        float *q = nullptr;
        rc = species_f32_ptr(h, s, &q);
        if (rc) return rc;
        rc = advect_species_exact(h, q);
        if (rc) return rc;
    }
    return 0;
}

int sync_fast_species(Handle *h)
{
    // The fused path keeps its own species table (and a second buffer per
    // species that it alternates with); refresh it after binding.
    std::vector<FastSpecies> v;
    for (int s = 0; s < h->nspecies; s++) {
        if (!h->sp[s].bound) continue;
        h->compact[s] = (int)v.size();
        v.push_back({h->sp[s].dev, h->sp[s].storage, h->sp[s].pow2});
    }
    int rc = fast_set_species(h->fast, v.data(), (int)v.size());
    if (rc) return set_err(h, rc, "%s", fast_error(h->fast));
    return 0;
}

}  // namespace

extern "C" {

int icar_adv_create(const icar_adv_grid *grid, const icar_adv_opts *opts, void **handle)
{
    if (!grid || !opts || !handle) return set_err(nullptr, ICAR_ADV_EINVAL, "NULL argument");
    *handle = nullptr;
    if (!grid_ok(grid))
        return set_err(nullptr, ICAR_ADV_ESHAPE,
                       "grid needs at least 3 cells in x, y and z and dx > 0 (got %d x %d x %d)",
                       grid->ime - grid->ims + 1, grid->jme - grid->jms + 1,
                       grid->kme - grid->kms + 1);
    if (opts->mpdata_order < 0 || opts->mpdata_order > 16)
        return set_err(nullptr, ICAR_ADV_EINVAL, "mpdata_order %d out of range [0,16]",
                       opts->mpdata_order);
    if (opts->kernel_variant < 0 || opts->kernel_variant > 1)
        return set_err(nullptr, ICAR_ADV_EINVAL, "kernel_variant %d unknown", opts->kernel_variant);
    Handle *h = new Handle();
    h->g = *grid;
    h->o = *opts;
    h->d = {grid->ime - grid->ims + 1, grid->jme - grid->jms + 1, grid->kme - grid->kms + 1};
    h->N = (size_t)h->d.nx * h->d.nz * h->d.ny;
    h->Nu = (size_t)(h->d.nx + 1) * h->d.nz * h->d.ny;
    h->Nv = (size_t)h->d.nx * h->d.nz * (h->d.ny + 1);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return set_err(nullptr, (int)e, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    const size_t fb = h->N * sizeof(float);
    float **bufs[] = {&h->sf.U, &h->sf.V, &h->sf.W, &h->sf.G, &h->sf.Gz, &h->sf.Wd};
    for (float **p : bufs)
        if (dalloc(h, (void **)p, fb)) {
            std::string m = h->err;
            free_all(h);
            return set_err(nullptr, ICAR_ADV_ENOMEM, "%s", m.c_str());
        }
    if (h->o.kernel_variant == 0) {
        float **work[] = {&h->q2, &h->u2, &h->v2, &h->w2, &h->u2l, &h->v2l, &h->w2l};
        for (float **p : work)
            if (dalloc(h, (void **)p, fb)) {
                std::string m = h->err;
                free_all(h);
                return set_err(nullptr, ICAR_ADV_ENOMEM, "%s", m.c_str());
            }
    } else {
        h->fast = fast_create(h->d);
        if (!h->fast) {
            free_all(h);
            return set_err(nullptr, ICAR_ADV_ENOMEM, "fused-path state allocation failed");
        }
    }
    const int nb = mass_partials_blocks(h->d);
    h->h_partial.assign((size_t)3 * nb, 0.0);
    if (dalloc(h, (void **)&h->d_partial, sizeof(double) * 3 * nb)) {
        std::string m = h->err;
        free_all(h);
        return set_err(nullptr, ICAR_ADV_ENOMEM, "%s", m.c_str());
    }
    *handle = h;
    return 0;
}

int icar_adv_destroy(void *handle)
{
    if (!handle) return 0;
    free_all(static_cast<Handle *>(handle));
    return 0;
}

int icar_adv_bind_field(void *handle, int field_id, float *host, const int lb[3], const int ub[3])
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (field_id < 0 || field_id >= ICAR_ADV_NFIELDS)
        return set_err(h, ICAR_ADV_EINVAL, "unknown field id %d", field_id);
    if (!host || !lb || !ub) return set_err(h, ICAR_ADV_EINVAL, "NULL array for field %d", field_id);
    const icar_adv_grid &g = h->g;
    int want_ub_i = g.ime, want_ub_j = g.jme;
    if (field_id == ICAR_ADV_FIELD_U || field_id == ICAR_ADV_FIELD_JACOBIAN_U) want_ub_i = g.ime + 1;
    if (field_id == ICAR_ADV_FIELD_V || field_id == ICAR_ADV_FIELD_JACOBIAN_V) want_ub_j = g.jme + 1;
    if (lb[0] != g.ims || ub[0] != want_ub_i || lb[1] != g.kms || ub[1] != g.kme ||
        lb[2] != g.jms || ub[2] != want_ub_j)
        return set_err(h, ICAR_ADV_ESHAPE,
                       "field %d bounds (%d:%d,%d:%d,%d:%d) do not match the grid "
                       "(%d:%d,%d:%d,%d:%d)",
                       field_id, lb[0], ub[0], lb[1], ub[1], lb[2], ub[2], g.ims, want_ub_i, g.kms,
                       g.kme, g.jms, want_ub_j);
    if (!h->dev_field[field_id]) {
        rc = dalloc(h, (void **)&h->dev_field[field_id], field_elems(h, field_id) * sizeof(float));
        if (rc) return rc;
    }
    h->host_field[field_id] = host;
    return 0;
}

int icar_adv_bind_species(void *handle, int idx, float *host, int storage, int pow2_exp)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (idx < 0 || idx >= ICAR_ADV_MAX_SPECIES)
        return set_err(h, ICAR_ADV_ERANGE, "species index %d outside [0,%d)", idx,
                       ICAR_ADV_MAX_SPECIES);
    if (!host) return set_err(h, ICAR_ADV_EINVAL, "NULL array for species %d", idx);
    if (storage != ICAR_ADV_STORE_F32 && storage != ICAR_ADV_STORE_BF16 &&
        storage != ICAR_ADV_STORE_F16)
        return set_err(h, ICAR_ADV_EINVAL, "unknown storage %d", storage);
    if (storage != ICAR_ADV_STORE_F32 && h->o.kernel_variant == 0)
        return set_err(h, ICAR_ADV_EINVAL,
                       "reduced-precision storage needs kernel_variant 1 (the exact kernels are "
                       "FP32 only)");
    if (pow2_exp < -126 || pow2_exp > 127)
        return set_err(h, ICAR_ADV_ERANGE, "pow2_exp %d out of range", pow2_exp);
    SpeciesSlot &s = h->sp[idx];
    const size_t elem = storage == ICAR_ADV_STORE_F32 ? 4 : 2;
    if (s.dev && s.storage != storage) {
        (void)hipFree(s.dev);
        s.dev = nullptr;
    }
    if (!s.dev) {
        rc = dalloc(h, &s.dev, h->N * elem);
        if (rc) return rc;
    }
    if (storage != ICAR_ADV_STORE_F32 && !h->stage) {
        rc = dalloc(h, (void **)&h->stage, h->N * sizeof(float));
        if (rc) return rc;
    }
    s.host = host;
    s.storage = storage;
    s.pow2 = storage == ICAR_ADV_STORE_F32 ? 0 : pow2_exp;
    s.bound = true;
    if (idx + 1 > h->nspecies) h->nspecies = idx + 1;
    if (h->fast) return sync_fast_species(h);
    return 0;
}

int icar_adv_upload(void *handle, unsigned long long mask)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        HIPCHK(h, hipMemcpyAsync(h->dev_field[f], h->host_field[f],
                                 field_elems(h, f) * sizeof(float), hipMemcpyHostToDevice,
                                 h->stream));
    }
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound) continue;
        if (sl.storage == ICAR_ADV_STORE_F32) {
            HIPCHK(h, hipMemcpyAsync(species_dev(h, s), sl.host, h->N * sizeof(float), hipMemcpyHostToDevice,
                                     h->stream));
        } else {
            HIPCHK(h, hipMemcpyAsync(h->stage, sl.host, h->N * sizeof(float),
                                     hipMemcpyHostToDevice, h->stream));
            timed(h, K_CONVERT, 6.0 * h->N,
                  [&] { launch_pack(h->stream, h->N, h->stage, species_dev(h, s), sl.storage, sl.pow2); });
        }
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_download(void *handle, unsigned long long mask)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        HIPCHK(h, hipMemcpyAsync(h->host_field[f], h->dev_field[f],
                                 field_elems(h, f) * sizeof(float), hipMemcpyDeviceToHost,
                                 h->stream));
    }
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound) continue;
        if (sl.storage == ICAR_ADV_STORE_F32) {
            HIPCHK(h, hipMemcpyAsync(sl.host, species_dev(h, s), h->N * sizeof(float), hipMemcpyDeviceToHost,
                                     h->stream));
        } else {
            timed(h, K_CONVERT, 6.0 * h->N,
                  [&] { launch_unpack(h->stream, h->N, species_dev(h, s), h->stage, sl.storage, sl.pow2); });
            HIPCHK(h, hipMemcpyAsync(sl.host, h->stage, h->N * sizeof(float),
                                     hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
        }
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_advect(void *handle, float dt)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    return do_advect(h, dt);
}

int icar_adv_mass_budget(void *handle, double *per_species)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!per_species) return set_err(h, ICAR_ADV_EINVAL, "NULL output");
    rc = require_bound(h);
    if (rc) return rc;
    const int nb = mass_partials_blocks(h->d);
    int out = 0;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!sl.bound) continue;
        const float *q = static_cast<const float *>(species_dev(h, s));
        if (sl.storage != ICAR_ADV_STORE_F32) {
            launch_unpack(h->stream, h->N, species_dev(h, s), h->stage, sl.storage, sl.pow2);
            q = h->stage;
        }
        timed(h, K_REDUCE, 4.0 * 4 * cells(h), [&] {
            launch_mass(h->stream, h->d, q, h->dev_field[ICAR_ADV_FIELD_JACOBIAN],
                        h->dev_field[ICAR_ADV_FIELD_DENSITY],
                        h->dev_field[ICAR_ADV_FIELD_ADVECTION_DZ], h->d_partial);
        });
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(h->h_partial.data(), h->d_partial, sizeof(double) * nb,
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        double sum = 0.0;
        for (int b = 0; b < nb; b++) sum += h->h_partial[b];
        per_species[out++] = sum;
    }
    return 0;
}

int icar_adv_max_courant(void *handle, double out[3])
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!out) return set_err(h, ICAR_ADV_EINVAL, "NULL output");
    if (!h->have_dt) return set_err(h, ICAR_ADV_EINVAL, "advect has not been called yet");
    const int nb = mass_partials_blocks(h->d);
    timed(h, K_REDUCE, 3.0 * 4 * cells(h),
          [&] { launch_courant_max(h->stream, h->d, h->sf.U, h->sf.V, h->sf.Wd, h->d_partial); });
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(h->h_partial.data(), h->d_partial, sizeof(double) * 3 * nb,
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out[0] = out[1] = out[2] = 0.0;
    for (int b = 0; b < nb; b++)
        for (int c = 0; c < 3; c++)
            out[c] = out[c] > h->h_partial[3 * b + c] ? out[c] : h->h_partial[3 * b + c];
    return 0;
}

const char *icar_adv_last_error(void *handle)
{
    if (!handle) return g_create_error.c_str();
    return static_cast<Handle *>(handle)->err.c_str();
}

int icar_adv_sync(void *handle)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_apply_bc(void *handle, int bc_x, int bc_y)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (bc_x < 0 || bc_x > 2 || (bc_y != 0 && bc_y != 2))
        return set_err(h, ICAR_ADV_EINVAL, "bc_x must be 0/1/2 and bc_y 0/2");
    if (bc_x == 0 && bc_y == 0) return 0;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!sl.bound) continue;
        timed(h, K_BC, 0.0, [&] {
            if (sl.storage == ICAR_ADV_STORE_F32)
                launch_bc(h->stream, h->d, static_cast<float *>(species_dev(h, s)), bc_x, bc_y);
            else
                launch_bc16(h->stream, h->d, species_dev(h, s), bc_x, bc_y);
        });
    }
    HIPCHK(h, hipGetLastError());
    return 0;
}

int icar_adv_run(void *handle, float dt, int nsteps, int bc_x, int bc_y)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (nsteps < 0) return set_err(h, ICAR_ADV_EINVAL, "nsteps < 0");
    for (int s = 0; s < nsteps; s++) {
        rc = do_advect(h, dt);
        if (rc) return rc;
        rc = icar_adv_apply_bc(handle, bc_x, bc_y);
        if (rc) return rc;
    }
    return 0;
}

int icar_adv_profile(void *handle, int enable)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drain_profile(h);
    if (enable) {
        for (int k = 0; k < K_COUNT; k++) {
            h->acc_ms[k] = 0;
            h->acc_n[k] = 0;
            h->acc_bytes[k] = 0;
        }
    }
    h->prof = enable != 0;
    return 0;
}

int icar_adv_kernel_count(void) { return K_COUNT; }

const char *icar_adv_kernel_name(int id)
{
    if (id < 0 || id >= K_COUNT) return "";
    return kernel_name(id);
}

int icar_adv_kernel_stats(void *handle, int id, double *total_ms, long long *launches,
                          double *alg_bytes)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (id < 0 || id >= K_COUNT) return set_err(h, ICAR_ADV_EINVAL, "unknown kernel id %d", id);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drain_profile(h);
    if (total_ms) *total_ms = h->acc_ms[id];
    if (launches) *launches = h->acc_n[id];
    if (alg_bytes) *alg_bytes = h->acc_bytes[id];
    return 0;
}

}  // extern "C"

namespace icaradv {
const char *kernel_name(int id)
{
    static const char *names[K_COUNT] = {"courant",     "upwind",     "pseudo_velocity",
                                         "fct",         "ring_bc",    "reduce",
                                         "copy",        "fused_first_pass", "fused_corrective_pass",
                                         "convert16",   "face_coefficients"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "";
}
}  // namespace icaradv
