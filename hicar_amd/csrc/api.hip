// C-ABI of the advection library: handle management, host<->device
// residency, the advect() orchestration and diagnostics.
//
// advect() follows src/physics/advection_driver.f90:59-126.  For adv=2:
// Courant numbers once (mpdata_compute_wind, adv_mpdata.f90:477-504), then
// each bound species through the order loop of mpdata_advect3d
// (adv_mpdata.f90:352-413).  For adv=1: the mass fluxes once
// (adv_std_compute_wind, advect.f90:444-492), then each species by the three
// RK3 calls of adv_std_advect3d (advection_driver.f90:98-110) or one Euler
// call (:117-118), the last call with the WRF limiter when flux_corr is 1.
// Species are independent inside advect, so the batched kernels advance
// many of them per launch with identical results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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
    float *fhi = nullptr, *dqdt = nullptr;  // lateral forcing (icar_adv_set_forcing)
};

struct TimedLaunch {
    int id;
    hipEvent_t a, b;
    double bytes;
};

// A memcpy split over a few persistent threads and the caller: the CPU side
// of the host transfers, where one thread copying pageable memory is what
// limits the runtime's own staged copies.
class CopyPool {
public:
    explicit CopyPool(int workers)
    {
        th_.reserve(workers);
        try {
            for (int i = 0; i < workers; i++) th_.emplace_back([this, i] { run(i + 1); });
        } catch (...) {  // the threads started so far end before the error leaves
            stop();
            throw;
        }
    }
    ~CopyPool() { stop(); }
    void copy(void *dst, const void *src, size_t n)
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            n_ = n;
            left_ = (int)th_.size();
            gen_++;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [this] { return left_ == 0; });
    }

private:
    void stop()
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
        th_.clear();
    }
    // part p of the workers + 1 contiguous pieces, cut at 64-byte boundaries
    void part(int p) const
    {
        const size_t parts = th_.size() + 1;
        const size_t a = p == 0 ? 0 : (n_ * p / parts) & ~(size_t)63;
        const size_t b = p + 1 == (int)parts ? n_ : (n_ * (p + 1) / parts) & ~(size_t)63;
        if (b > a) memcpy(dst_ + a, src_ + a, b - a);
    }
    void run(int p)
    {
        unsigned seen = 0;
        std::unique_lock<std::mutex> l(mu_);
        for (;;) {
            cv_.wait(l, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            l.unlock();
            part(p);
            l.lock();
            if (--left_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    int left_ = 0;
    unsigned gen_ = 0;
    bool quit_ = false;
};

// one host<->device copy of a transfer
struct XferItem {
    char *host;
    char *dev;
    size_t bytes;
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
    float *d_dt = nullptr;  // dz_levels (nz) then the dt partial maxima (4 per block)
    float *d_filter = nullptr;  // relaxation filter (nx * ny), icar_adv_set_relax_filter
    std::string err;
    FastState *fast = nullptr;
    FastBatch *batch = nullptr;  // this domain as a batch of one (fused path)
    // counts species bindings: each may reallocate device buffers and the
    // fused path's tables, which recorded graphs and ensembles point at
    unsigned bind_gen = 0;
    int compensated = 0;  // icar_adv_set_compensation
    // graph capture of icar_adv_run
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    float g_dt = 0.0f;
    int g_steps = -1, g_bcx = -1, g_bcy = -1, g_par0 = -1, g_par1 = -1;
    // standard scheme (adv=1): mass fluxes, divisors, and per species the
    // buffer the next state is written to (swapped with the species' own
    // after each step: std_par counts the swaps mod 2), the second RK3 stage
    // and the limiter's dumb_q and (scale_in, scale_out) pairs
    bool is_std = false;
    float *Um = nullptr, *Vm = nullptr, *Wm = nullptr, *Rj = nullptr, *Rdj = nullptr;
    float *std_alt[ICAR_ADV_MAX_SPECIES] = {}, *std_b[ICAR_ADV_MAX_SPECIES] = {};
    float *std_dq[ICAR_ADV_MAX_SPECIES] = {};
    float2 *std_sc[ICAR_ADV_MAX_SPECIES] = {};
    int std_par = 0;
    // output / restart snapshot: FP32 device staging, pinned host staging,
    // a copy stream, and what to copy out when it is complete
    float *snap_dev = nullptr, *snap_host = nullptr;
    size_t snap_cap = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    bool snap_pending = false;
    unsigned long long snap_mask = 0;  // the upload/download bits of the pending snapshot
    struct SnapItem {
        float *host;
        size_t off, n;
    };
    std::vector<SnapItem> snap_items;
    // upload / download through pinned staging: two chunks (the DMA of one
    // overlaps the CPU copy of the other) and their events
    char *xfer_pin[2] = {};
    hipEvent_t xfer_ev[2] = {};
    bool xfer_failed = false;  // pinned staging unavailable: the runtime's copies
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

template <class T>
hipEvent_t pool_event(T *h)
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
template <class T, class F>
void timed(T *h, int id, double bytes, F &&launch)
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

template <class T>
void drain_profile(T *h)
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
    for (auto &s : h->sp) {
        (void)hipFree(s.dev);
        (void)hipFree(s.fhi);
        (void)hipFree(s.dqdt);
    }
    float *sbufs[] = {h->Um, h->Vm, h->Wm, h->Rj, h->Rdj};
    for (float *p : sbufs) (void)hipFree(p);
    for (int s = 0; s < ICAR_ADV_MAX_SPECIES; s++) {
        void *per[] = {h->std_alt[s], h->std_b[s], h->std_dq[s], h->std_sc[s]};
        for (void *p : per) (void)hipFree(p);
    }
    (void)hipFree(h->d_filter);
    (void)hipFree(h->d_partial);
    (void)hipFree(h->d_dt);
    if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
    (void)hipFree(h->snap_dev);
    if (h->snap_host) (void)hipHostFree(h->snap_host);
    if (h->snap_ready) (void)hipEventDestroy(h->snap_ready);
    if (h->snap_done) (void)hipEventDestroy(h->snap_done);
    if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
    for (int b = 0; b < 2; b++) {
        if (h->xfer_pin[b]) (void)hipHostFree(h->xfer_pin[b]);
        if (h->xfer_ev[b]) (void)hipEventDestroy(h->xfer_ev[b]);
    }
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    fast_batch_destroy(h->batch);
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

int ensure_batch(Handle *h)
{
    if (h->batch) return 0;
    FastDomain fd{h->fast, domain_fields(h), h->sf, h->g.dx, h->o.advect_density,
                  h->compensated};
    std::string e;
    h->batch = fast_batch_create(&fd, 1, h->o.mpdata_order, h->o.mpdata_fct, &e);
    if (!h->batch) return set_err(h, ICAR_ADV_EINVAL, "%s", e.c_str());
    return 0;
}

template <class T>
TimedFn timed_fn(T *h)
{
    return [h](int id, double bytes, const std::function<void()> &fn) { timed(h, id, bytes, fn); };
}

// test_divergence (adv_mpdata.f90:445-475) on the Courant numbers of the last
// MPDATA step: the number of interior cells with |div| > 1e-3 and the largest
// |div| (an FP32 value, exact in the double).  Synchronises.
int divergence_of_step(Handle *h, double *cnt_out, double *max_out)
{
    const int nb = mass_partials_blocks(h->d);
    launch_divergence(h->stream, h->d, h->sf.U, h->sf.V, h->sf.W,
                      h->dev_field[ICAR_ADV_FIELD_ADVECTION_DZ], h->d_partial);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(h->h_partial.data(), h->d_partial, sizeof(double) * 2 * nb,
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double cnt = 0, mx = 0;
    for (int b = 0; b < nb; b++) {
        cnt += h->h_partial[2 * b];
        mx = mx > h->h_partial[2 * b + 1] ? mx : h->h_partial[2 * b + 1];
    }
    *cnt_out = cnt;
    *max_out = mx;
    return 0;
}

// test_divergence (adv_mpdata.f90:500-502): the reference prints a warning
int divergence_check(Handle *h)
{
    double cnt = 0, mx = 0;
    const int rc = divergence_of_step(h, &cnt, &mx);
    if (rc) return rc;
    if (cnt > 0)
        fprintf(stderr,
                "icar_adv: winds are not balanced on entry to advect: %.0f cells with "
                "|div| > 1e-3 (max %g)\n",
                cnt, mx);
    return 0;
}

// adv=1: the standard scheme for every bound species, in batches of up to
// STD_MAX_BATCH species per launch.
int advect_std(Handle *h, float dt)
{
    const DomainFields f = domain_fields(h);
    const double n = cells(h);
    timed(h, K_STD_WIND, 14.0 * 4 * n, [&] {  // 9 inputs, 5 outputs
        launch_std_wind(h->stream, h->d, h->g.dx, dt, h->o.advect_density, f, h->Um, h->Vm,
                        h->Wm, h->Rj, h->Rdj);
    });
    StdGeom g;
    g.d = h->d;
    g.its = h->g.its - h->g.ims;
    g.ite = h->g.ite - h->g.ims;
    g.jts = h->g.jts - h->g.jms;
    g.jte = h->g.jte - h->g.jms;
    g.horder = h->o.h_order;
    g.vorder = h->o.v_order;
    g.Um = h->Um;
    g.Vm = h->Vm;
    g.Wm = h->Wm;
    g.Rj = h->Rj;
    g.Rdj = h->Rdj;
    const bool fc = h->o.flux_corr == 1, rk3 = h->o.rk3 != 0;
    std::vector<int> idx;
    for (int s = 0; s < h->nspecies; s++)
        if (h->sp[s].bound) idx.push_back(s);
    for (size_t b0 = 0; b0 < idx.size(); b0 += STD_MAX_BATCH) {
        StdBatch b{};
        b.n = (int)std::min(idx.size() - b0, (size_t)STD_MAX_BATCH);
        const double ns = (double)b.n;
        auto fill = [&](int which_in) {
            for (int m = 0; m < b.n; m++) {
                const int s = idx[b0 + m];
                float *home = static_cast<float *>(h->sp[s].dev);
                b.qold[m] = home;
                b.qin[m] = which_in == 0 ? home : which_in == 1 ? h->std_alt[s] : h->std_b[s];
                b.dq[m] = h->std_dq[s];
                b.sc[m] = h->std_sc[s];
            }
        };
        auto outs = [&](int which_out) {
            for (int m = 0; m < b.n; m++) {
                const int s = idx[b0 + m];
                b.out[m] = which_out == 1 ? h->std_alt[s] : h->std_b[s];
            }
        };
        // algorithmic bytes: winds and divisors once, then per species the
        // arrays the kernel reads or writes once each (the scale pair counts 2)
        const double wb = 5.0 * 4 * n;
        // the cells outside the tile keep the base state in every call's
        // output (qfluxes = qold, advect.f90:298), dumb_q outside its ring
        fill(0);
        outs(1);
        timed(h, K_STD_RING, 0.0, [&] { launch_std_ring(h->stream, g, b, fc ? 1 : 0); });
        if (rk3) {
            outs(2);
            timed(h, K_STD_RING, 0.0, [&] { launch_std_ring(h->stream, g, b, 0); });
        }
        if (rk3) {  // advection_driver.f90:103-110
            fill(0);
            outs(1);
            timed(h, K_STD_STAGE, wb + ns * (fc ? 3 : 2) * 4 * n,
                  [&] { launch_std_stage(h->stream, g, 0.333f, fc ? 1 : 0, b); });
            fill(1);
            outs(2);
            timed(h, K_STD_STAGE, wb + ns * 3 * 4 * n,
                  [&] { launch_std_stage(h->stream, g, 0.5f, 0, b); });
            fill(2);
            outs(1);
        } else {  // :118, the species is both qfluxes and qold
            fill(0);
            outs(1);
            if (fc)
                timed(h, K_STD_STAGE, wb + ns * 2 * 4 * n,
                      [&] { launch_std_stage(h->stream, g, 1.0f, 2, b); });
        }
        if (fc) {
            timed(h, K_STD_SCALE, wb + ns * 5 * 4 * n, [&] { launch_std_scale(h->stream, g, b); });
            timed(h, K_STD_FINAL, wb + ns * 6 * 4 * n, [&] { launch_std_final(h->stream, g, b); });
        } else {
            timed(h, K_STD_STAGE, wb + ns * (rk3 ? 3 : 2) * 4 * n,
                  [&] { launch_std_stage(h->stream, g, 1.0f, 0, b); });
        }
    }
    HIPCHK(h, hipGetLastError());
    // the new states are in the alternate buffers: make them the species'
    for (int s : idx) {
        void *cur = h->sp[s].dev;
        h->sp[s].dev = h->std_alt[s];
        h->std_alt[s] = static_cast<float *>(cur);
    }
    h->std_par ^= 1;
    return 0;
}

// Which of its two buffers each species lives in: the fused MPDATA path and
// the standard scheme alternate buffers every step.  A recorded graph holds
// the buffer addresses, so it replays only from the parity it was recorded at.
int parity(Handle *h)
{
    if (h->fast) return fast_parity(h->fast);
    return h->is_std ? h->std_par : 0;
}

void set_parity(Handle *h, int p)
{
    if (h->fast) {
        fast_set_parity(h->fast, p);
    } else if (h->is_std && h->std_par != p) {
        for (int s = 0; s < h->nspecies; s++) {
            if (!h->sp[s].bound) continue;
            void *cur = h->sp[s].dev;
            h->sp[s].dev = h->std_alt[s];
            h->std_alt[s] = static_cast<float *>(cur);
        }
        h->std_par = p;
    }
}

int do_advect(Handle *h, float dt)
{
    int rc = require_bound(h);
    if (rc) return rc;
    h->last_dt = dt;
    h->have_dt = true;
    if (h->is_std) return advect_std(h, dt);
    const DomainFields f = domain_fields(h);
    const double n = cells(h);
    if (h->o.kernel_variant == 1 && !h->o.rk3 && h->nspecies > 0 && h->o.mpdata_order >= 1) {
        // fused path: Courant numbers, coefficients and every pass of every
        // species in one launch per pass
        rc = ensure_batch(h);
        if (rc) return rc;
        std::string e;
        rc = fast_batch_advect(h->batch, h->stream, &dt, timed_fn(h), &e);
        if (rc) return set_err(h, rc, "%s", e.c_str());
        HIPCHK(h, hipGetLastError());
        if (h->o.debug) return divergence_check(h);
        return 0;
    }
    timed(h, K_COURANT, (9 + 6) * 4 * n, [&] {
        launch_courant(h->stream, h->d, h->g.dx, dt, h->o.advect_density, f, h->sf);
    });
    HIPCHK(h, hipGetLastError());
    if (h->o.debug) {
        rc = divergence_check(h);
        if (rc) return rc;
    }
    if (h->o.rk3) return 0;  // advection_driver.f90:112-115
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
    fast_batch_destroy(h->batch);  // its tables point at the old buffers
    h->batch = nullptr;
    int rc = fast_set_species(h->fast, v.data(), (int)v.size(), h->stream);
    if (rc) return set_err(h, rc, "%s", fast_error(h->fast));
    return 0;
}

// Ring conditions for every species of h, launched on stream s and timed
// by owner (the handle itself, or an ensemble).
template <class T>
int apply_bc_on(Handle *h, T *owner, hipStream_t stream, int bc_x, int bc_y)
{
    if (bc_x < 0 || bc_x > 2 || (bc_y != 0 && bc_y != 2))
        return set_err(h, ICAR_ADV_EINVAL, "bc_x must be 0/1/2 and bc_y 0/2");
    if (bc_x == 0 && bc_y == 0) return 0;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!sl.bound) continue;
        timed(owner, K_BC, 0.0, [&] {
            if (sl.storage == ICAR_ADV_STORE_F32)
                launch_bc(stream, h->d, static_cast<float *>(species_dev(h, s)), bc_x, bc_y);
            else
                launch_bc16(stream, h->d, species_dev(h, s), bc_x, bc_y);
        });
    }
    HIPCHK(h, hipGetLastError());
    return 0;
}

// The strip a side of the tile sends (inside the tile, as wide as the halo
// on that side) and the halo it fills on the neighbour's opposite side,
// both over the tile's range along the side (exchangeable_obj.f90:240-392,
// domain_obj.f90:573-650: corners are not exchanged).  0-based offsets.
struct Box {
    int i0, ni, j0, nj;
};

bool side_boxes(const Handle *h, int side, Box *send, Box *recv)
{
    const icar_adv_grid &g = h->g;
    const int its = g.its - g.ims, ite = g.ite - g.ims, jts = g.jts - g.jms, jte = g.jte - g.jms;
    const int nx = h->d.nx, ny = h->d.ny, ti = ite - its + 1, tj = jte - jts + 1;
    int w;
    switch (side) {
    case ICAR_ADV_NORTH:  // rows jte-w+1..jte -> the neighbour's rows jms..jts-1
        w = ny - 1 - jte;
        *send = {its, ti, jte - w + 1, w};
        *recv = {its, ti, jte + 1, w};
        break;
    case ICAR_ADV_SOUTH:
        w = jts;
        *send = {its, ti, jts, w};
        *recv = {its, ti, 0, w};
        break;
    case ICAR_ADV_EAST:
        w = nx - 1 - ite;
        *send = {ite - w + 1, w, jts, tj};
        *recv = {ite + 1, w, jts, tj};
        break;
    case ICAR_ADV_WEST:
        w = its;
        *send = {its, w, jts, tj};
        *recv = {0, w, jts, tj};
        break;
    default:
        return false;
    }
    return w >= 1 && w <= ti && w <= tj;
}

int opposite(int side)
{
    return side == ICAR_ADV_NORTH ? ICAR_ADV_SOUTH
           : side == ICAR_ADV_SOUTH ? ICAR_ADV_NORTH
           : side == ICAR_ADV_EAST  ? ICAR_ADV_WEST
                                    : ICAR_ADV_EAST;
}

// bytes of box b for every bound species of h; fills the species table
size_t strip_table(Handle *h, const Box &b, HaloStrips *hs, bool as_source)
{
    size_t bytes = 0;
    hs->n = 0;
    for (int s = 0; s < h->nspecies; s++) {
        if (!h->sp[s].bound) continue;
        const int m = hs->n++;
        const int es = h->sp[s].storage == ICAR_ADV_STORE_F32 ? 4 : 2;
        if (as_source) hs->a[m] = species_dev(h, s);
        else hs->b[m] = species_dev(h, s);
        hs->esize[m] = es;
        hs->buf_off[m] = (long long)bytes;
        bytes += (size_t)es * b.ni * b.nj * h->d.nz;
        bytes = (bytes + 15) & ~(size_t)15;
    }
    return bytes;
}

// Host transfers of at least XFER_MIN bytes go through the handle's pinned
// staging in XFER_CHUNK pieces: the runtime stages pageable copies through
// small buffers with one CPU thread; here the DMA of one chunk overlaps the
// copy of the previous one by several threads.  Smaller transfers, or a
// process where pinned memory cannot be had, use the runtime's copies.
constexpr size_t XFER_CHUNK = 32u << 20, XFER_MIN = 8u << 20;

// copy-pool workers besides the caller: ICAR_ADV_COPY_THREADS, default 7
// (0: every copy on the calling thread)
int xfer_threads()
{
    static const int n = [] {
        const char *e = std::getenv("ICAR_ADV_COPY_THREADS");
        const int v = e ? std::atoi(e) : 7;
        return std::max(0, std::min(v, 64));
    }();
    return n;
}

// The copy pool, one for the process: host copies are bound by host memory
// bandwidth, so the handles share it, one copy at a time (a copy holds it
// while it runs).  Created on first use; when its threads cannot be created
// (no exception leaves the C-ABI) the copies run on the calling thread.
struct SharedPool {
    std::mutex use;
    std::unique_ptr<CopyPool> pool;
    bool failed = false;
};

SharedPool &shared_pool()
{
    static SharedPool p;
    return p;
}

void host_copy(Handle *, void *dst, const void *src, size_t n)
{
    SharedPool &sp = shared_pool();
    std::lock_guard<std::mutex> l(sp.use);
    if (!sp.pool && !sp.failed && xfer_threads() > 0) {
        try {
            sp.pool.reset(new CopyPool(xfer_threads()));
        } catch (...) {
            sp.failed = true;
        }
    }
    if (sp.pool)
        sp.pool->copy(dst, src, n);
    else
        memcpy(dst, src, n);
}

bool xfer_ready(Handle *h)
{
    if (h->xfer_failed) return false;
    if (h->xfer_pin[0]) return true;
    for (int b = 0; b < 2; b++) {
        // non-coherent: cacheable for the CPU side of the copies
        if (hipHostMalloc((void **)&h->xfer_pin[b], XFER_CHUNK, hipHostMallocNonCoherent) !=
                hipSuccess ||
            hipEventCreateWithFlags(&h->xfer_ev[b], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            h->xfer_failed = true;
            return false;
        }
    }
    return true;
}

size_t xfer_bytes(const std::vector<XferItem> &items)
{
    size_t n = 0;
    for (const auto &it : items) n += it.bytes;
    return n;
}

// device -> host, in stream order after the work enqueued so far; returns
// once the host arrays hold the data
int xfer_down(Handle *h, const std::vector<XferItem> &items)
{
    if (xfer_bytes(items) < XFER_MIN || !xfer_ready(h)) {
        for (const auto &it : items)
            HIPCHK(h, hipMemcpyAsync(it.host, it.dev, it.bytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return 0;
    }
    struct Piece {
        char *host;
        const char *dev;
        size_t n;
    };
    std::vector<Piece> ps;
    for (const auto &it : items)
        for (size_t o = 0; o < it.bytes; o += XFER_CHUNK)
            ps.push_back({it.host + o, it.dev + o, std::min(XFER_CHUNK, it.bytes - o)});
    for (size_t i = 0; i <= ps.size(); i++) {
        if (i < ps.size()) {  // chunk i into buffer i%2 (its previous contents are copied out)
            HIPCHK(h, hipMemcpyAsync(h->xfer_pin[i & 1], ps[i].dev, ps[i].n, hipMemcpyDeviceToHost,
                                     h->stream));
            HIPCHK(h, hipEventRecord(h->xfer_ev[i & 1], h->stream));
        }
        if (i > 0) {  // chunk i-1 out of its buffer while chunk i moves
            const size_t b = (i - 1) & 1;
            HIPCHK(h, hipEventSynchronize(h->xfer_ev[b]));
            host_copy(h, ps[i - 1].host, h->xfer_pin[b], ps[i - 1].n);
        }
    }
    return 0;
}

// host -> device, enqueued on the handle's stream; the host arrays may be
// changed once this returns (their contents are in the staging or on the
// device), the device copies are complete in stream order
int xfer_up(Handle *h, const std::vector<XferItem> &items)
{
    if (xfer_bytes(items) < XFER_MIN || !xfer_ready(h)) {
        for (const auto &it : items)
            HIPCHK(h, hipMemcpyAsync(it.dev, it.host, it.bytes, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return 0;
    }
    size_t i = 0;
    for (const auto &it : items)
        for (size_t o = 0; o < it.bytes; o += XFER_CHUNK, i++) {
            const size_t b = i & 1, n = std::min(XFER_CHUNK, it.bytes - o);
            // the buffer's last DMA has left: chunk i-2's, or for i < 2 one a
            // previous call enqueued (complete unless that call failed part way;
            // an event never recorded returns at once)
            HIPCHK(h, hipEventSynchronize(h->xfer_ev[b]));
            host_copy(h, h->xfer_pin[b], it.host + o, n);
            HIPCHK(h, hipMemcpyAsync(it.dev + o, h->xfer_pin[b], n, hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipEventRecord(h->xfer_ev[b], h->stream));
        }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// Completes a pending snapshot into the host arrays.
int finish_snapshot(Handle *h)
{
    if (!h->snap_pending) return 0;
    h->snap_pending = false;
    HIPCHK(h, hipEventSynchronize(h->snap_done));
    for (const auto &it : h->snap_items) {
        if (it.n * sizeof(float) >= XFER_MIN)
            host_copy(h, it.host, h->snap_host + it.off, it.n * sizeof(float));
        else
            memcpy(it.host, h->snap_host + it.off, it.n * sizeof(float));
    }
    h->snap_items.clear();
    return 0;
}

// Options and bounds the standard scheme needs: orders 1/3/5, flux_corr 0/1,
// and a halo wide enough for flux3's stencil (plus one cell for the limiter's
// extended tile), as ICAR's grid gives it (domain_obj.f90:2979-2989).
int std_check(const icar_adv_grid *g, const icar_adv_opts *o)
{
    auto odd = [](int v) { return v == 1 || v == 3 || v == 5; };
    if (!odd(o->h_order) || !odd(o->v_order))
        return set_err(nullptr, ICAR_ADV_EINVAL, "h_order %d / v_order %d: only 1, 3 and 5",
                       o->h_order, o->v_order);
    if (o->flux_corr != 0 && o->flux_corr != 1)
        return set_err(nullptr, ICAR_ADV_EINVAL, "flux_corr %d unknown (0 none, 1 WRF monotonic)",
                       o->flux_corr);
    const int need = (o->h_order + 1) / 2 + (o->flux_corr == 1 ? 1 : 0);
    const int hw = std::min(g->its - g->ims, g->ime - g->ite);
    const int hs = std::min(g->jts - g->jms, g->jme - g->jte);
    if (g->ite < g->its || g->jte < g->jts || hw < need || hs < need)
        return set_err(nullptr, ICAR_ADV_ESHAPE,
                       "h_order %d%s needs a halo of %d cells around the tile (have %d in x, %d "
                       "in y)",
                       o->h_order, o->flux_corr == 1 ? " with the limiter" : "", need, hw, hs);
    if (o->v_order == 5 && g->kme - g->kms + 1 < 4)
        return set_err(nullptr, ICAR_ADV_ESHAPE, "v_order 5 needs at least 4 levels");
    // the kernels form the byte offsets of float arrays in 32 bits
    const long long cells = (long long)(g->ime - g->ims + 2) * (g->jme - g->jms + 2) *
                            (g->kme - g->kms + 1);
    if (cells >= (1LL << 30))
        return set_err(nullptr, ICAR_ADV_ERANGE,
                       "the standard scheme takes at most 2^30 cells per domain (%lld)", cells);
    return 0;
}

}  // namespace

extern "C" {

int icar_adv_create(const icar_adv_grid *grid, const icar_adv_opts *opts, void **handle)
{
    int rc_create = 0;
    if (!grid || !opts || !handle) return set_err(nullptr, ICAR_ADV_EINVAL, "NULL argument");
    *handle = nullptr;
    if (!grid_ok(grid))
        return set_err(nullptr, ICAR_ADV_ESHAPE,
                       "grid needs at least 3 cells in x, y and z and dx > 0 (got %d x %d x %d)",
                       grid->ime - grid->ims + 1, grid->jme - grid->jms + 1,
                       grid->kme - grid->kms + 1);
    // the tile bounds inside the memory bounds: the halo strips, the standard
    // scheme and the dt bound address the tile.  An empty tile (its > ite, as
    // a 3-cell-wide domain with a halo of 2 has) is allowed: MPDATA does not
    // use it, and the calls that do refuse it.
    if (grid->its < grid->ims || grid->its > grid->ime || grid->ite < grid->ims ||
        grid->ite > grid->ime || grid->jts < grid->jms || grid->jts > grid->jme ||
        grid->jte < grid->jms || grid->jte > grid->jme)
        return set_err(nullptr, ICAR_ADV_ESHAPE,
                       "tile %d..%d x %d..%d must lie inside the memory bounds %d..%d x %d..%d",
                       grid->its, grid->ite, grid->jts, grid->jte, grid->ims, grid->ime, grid->jms,
                       grid->jme);
    if (opts->mpdata_order < 0 || opts->mpdata_order > 16)
        return set_err(nullptr, ICAR_ADV_EINVAL, "mpdata_order %d out of range [0,16]",
                       opts->mpdata_order);
    if (opts->kernel_variant < 0 || opts->kernel_variant > 1)
        return set_err(nullptr, ICAR_ADV_EINVAL, "kernel_variant %d unknown", opts->kernel_variant);
    const int adv = opts->advection == 0 ? ICAR_ADV_MPDATA : opts->advection;
    if (adv != ICAR_ADV_STD && adv != ICAR_ADV_MPDATA)
        return set_err(nullptr, ICAR_ADV_EINVAL, "advection %d unknown (1 standard, 2 MPDATA)",
                       opts->advection);
    if (adv == ICAR_ADV_STD) {
        rc_create = std_check(grid, opts);
        if (rc_create) return rc_create;
    }
    Handle *h = new Handle();
    h->is_std = adv == ICAR_ADV_STD;
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
    // MPDATA's Courant numbers and products; the standard scheme keeps its own
    // (U_m, V_m, W_m and the divisors below) and leaves these null
    float **bufs[] = {&h->sf.U, &h->sf.V, &h->sf.W, &h->sf.G, &h->sf.Gz, &h->sf.Wd};
    for (float **p : bufs)
        if (!h->is_std && dalloc(h, (void **)p, fb)) {
            std::string m = h->err;
            free_all(h);
            return set_err(nullptr, ICAR_ADV_ENOMEM, "%s", m.c_str());
        }
    if (h->is_std) {  // mass fluxes and divisors of the standard scheme
        const size_t sz[] = {h->Nu * sizeof(float), h->Nv * sizeof(float), fb, fb, fb};
        float **sb[] = {&h->Um, &h->Vm, &h->Wm, &h->Rj, &h->Rdj};
        for (int q = 0; q < 5; q++)
            if (dalloc(h, (void **)sb[q], sz[q])) {
                std::string m = h->err;
                free_all(h);
                return set_err(nullptr, ICAR_ADV_ENOMEM, "%s", m.c_str());
            }
    } else if (h->o.kernel_variant == 0) {
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
    Handle *h = static_cast<Handle *>(handle);
    const int rc = finish_snapshot(h);  // a pending snapshot reaches the host arrays first
    free_all(h);
    return rc;
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
    if (storage < ICAR_ADV_STORE_F32 || storage > ICAR_ADV_STORE_F16_SR)
        return set_err(h, ICAR_ADV_EINVAL, "unknown storage %d", storage);
    if (storage != ICAR_ADV_STORE_F32 && h->is_std)
        return set_err(h, ICAR_ADV_EINVAL, "the standard scheme keeps FP32 storage");
    if (storage != ICAR_ADV_STORE_F32 && h->o.kernel_variant == 0)
        return set_err(h, ICAR_ADV_EINVAL,
                       "reduced-precision storage needs kernel_variant 1 (the exact kernels are "
                       "FP32 only)");
    if (pow2_exp < -126 || pow2_exp > 127)
        return set_err(h, ICAR_ADV_ERANGE, "pow2_exp %d out of range", pow2_exp);
    SpeciesSlot &s = h->sp[idx];
    const size_t elem = storage == ICAR_ADV_STORE_F32 ? 4 : 2;
    // a recorded run holds the old buffers and species list: record anew
    h->bind_gen++;
    if (h->gexec) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    if (s.dev && s.storage != storage) {
        (void)hipFree(s.dev);
        s.dev = nullptr;
    }
    if (!s.dev) {
        rc = dalloc(h, &s.dev, h->N * elem + 64);  // the fused kernels read 16-bit data in aligned pairs
        if (rc) return rc;
    }
    if (storage != ICAR_ADV_STORE_F32 && !h->stage) {
        rc = dalloc(h, (void **)&h->stage, h->N * sizeof(float));
        if (rc) return rc;
    }
    if (h->is_std) {  // next-state buffer, second RK3 stage, limiter work arrays
        void **per[] = {(void **)&h->std_alt[idx], (void **)&h->std_b[idx],
                        (void **)&h->std_dq[idx], (void **)&h->std_sc[idx]};
        const bool need[] = {true, h->o.rk3 != 0, h->o.flux_corr == 1, h->o.flux_corr == 1};
        const size_t elem[] = {4, 4, 4, 8};
        for (int q = 0; q < 4; q++)
            if (need[q] && !*per[q]) {
                rc = dalloc(h, per[q], h->N * elem[q] + 64);
                if (rc) return rc;
            }
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
    // A pending snapshot has yet to write its host arrays.  Completing it here
    // would overwrite what the caller wrote into them since, and completing it
    // later would overwrite them after this upload read them: refuse instead.
    if (h->snap_pending && (mask & h->snap_mask))
        return set_err(h, ICAR_ADV_EINVAL,
                       "a snapshot of these arrays is pending: icar_adv_snapshot_wait before "
                       "uploading them");
    // FP32 arrays in one pipelined transfer; each 16-bit species through the
    // FP32 staging and the pack kernel (stream order keeps the staging's uses
    // apart)
    std::vector<XferItem> items;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        items.push_back({(char *)h->host_field[f], (char *)h->dev_field[f],
                         field_elems(h, f) * sizeof(float)});
    }
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound) continue;
        if (sl.storage == ICAR_ADV_STORE_F32)
            items.push_back({(char *)sl.host, (char *)species_dev(h, s), h->N * sizeof(float)});
    }
    rc = xfer_up(h, items);
    if (rc) return rc;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound || sl.storage == ICAR_ADV_STORE_F32)
            continue;
        rc = xfer_up(h, {{(char *)sl.host, (char *)h->stage, h->N * sizeof(float)}});
        if (rc) return rc;
        timed(h, K_CONVERT, 6.0 * h->N,
              [&] { launch_pack(h->stream, h->N, h->stage, species_dev(h, s), sl.storage, sl.pow2); });
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
    // a pending snapshot writes the host arrays first, so that the newer
    // values of this download are the ones that stay
    rc = finish_snapshot(h);
    if (rc) return rc;
    std::vector<XferItem> items;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        items.push_back({(char *)h->host_field[f], (char *)h->dev_field[f],
                         field_elems(h, f) * sizeof(float)});
    }
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound) continue;
        if (sl.storage == ICAR_ADV_STORE_F32)
            items.push_back({(char *)sl.host, (char *)species_dev(h, s), h->N * sizeof(float)});
    }
    rc = xfer_down(h, items);
    if (rc) return rc;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound || sl.storage == ICAR_ADV_STORE_F32)
            continue;
        timed(h, K_CONVERT, 6.0 * h->N,
              [&] { launch_unpack(h->stream, h->N, species_dev(h, s), h->stage, sl.storage, sl.pow2); });
        HIPCHK(h, hipGetLastError());
        rc = xfer_down(h, {{(char *)sl.host, (char *)h->stage, h->N * sizeof(float)}});
        if (rc) return rc;
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
    timed(h, K_REDUCE, (h->is_std ? 4.0 : 3.0) * 4 * cells(h), [&] {
        if (h->is_std)  // the standard scheme's U_m, V_m, W_m of its last step
            launch_courant_max(h->stream, h->d, h->Um, h->d.nx + 1, h->Vm, h->Wm,
                               h->dev_field[ICAR_ADV_FIELD_ADVECTION_DZ], h->d_partial);
        else
            launch_courant_max(h->stream, h->d, h->sf.U, h->d.nx, h->sf.V, h->sf.Wd, nullptr,
                               h->d_partial);
    });
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

int icar_adv_set_compensation(void *handle, int on)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!h->fast)
        return set_err(h, ICAR_ADV_EINVAL,
                       "compensated sums are an option of the fused MPDATA kernels "
                       "(kernel_variant 1)");
    h->compensated = on ? 1 : 0;
    // the batch's tables and a recorded run hold the old setting; ensembles
    // built over this handle refuse to run, as after a binding
    h->bind_gen++;
    if (h->gexec) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    fast_batch_destroy(h->batch);
    h->batch = nullptr;
    return 0;
}

int icar_adv_divergence(void *handle, double out[2])
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!out) return set_err(h, ICAR_ADV_EINVAL, "NULL output");
    if (h->is_std)
        return set_err(h, ICAR_ADV_EINVAL,
                       "the divergence check belongs to MPDATA (the standard scheme has none)");
    if (!h->have_dt) return set_err(h, ICAR_ADV_EINVAL, "advect has not been called yet");
    return divergence_of_step(h, &out[0], &out[1]);
}

int icar_adv_balance_w(void *handle)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    rc = require_bound(h);
    if (rc) return rc;
    // columns its-1..ite+1, jts-1..jte+1 clipped to the domain (wind.f90:113-122);
    // at the memory bounds, which are the domain's at its edges
    const icar_adv_grid &g = h->g;
    const int i_s = std::max(g.its - 1, g.ims) - g.ims, i_e = std::min(g.ite + 1, g.ime) - g.ims;
    const int j_s = std::max(g.jts - 1, g.jms) - g.jms, j_e = std::min(g.jte + 1, g.jme) - g.jms;
    DomainFields f = domain_fields(h);
    // reads u, jacobian_u, v, jacobian_v, density, dz, jacobian_w; writes w
    timed(h, K_BALANCE, 4.0 * (2.0 * h->Nu + 2.0 * h->Nv + 4.0 * h->N), [&] {
        launch_balance_w(h->stream, h->d, i_s, i_e, j_s, j_e, g.dx, h->o.advect_density, f,
                         h->dev_field[ICAR_ADV_FIELD_W]);
    });
    HIPCHK(h, hipGetLastError());
    if (h->o.debug) HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_compute_dt(void *handle, float cfl, int cfl_strictness, const float *dz_levels,
                        float *dt)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!dz_levels || !dt) return set_err(h, ICAR_ADV_EINVAL, "NULL dz_levels or dt");
    rc = require_bound(h);
    if (rc) return rc;
    const icar_adv_grid &g = h->g;
    const Dims &d = h->d;
    const int its = std::max(g.its, g.ims) - g.ims, ite = std::min(g.ite, g.ime) - g.ims;
    const int jts = std::max(g.jts, g.jms) - g.jms, jte = std::min(g.jte, g.jme) - g.jms;
    if (its > ite || jts > jte)
        return set_err(h, ICAR_ADV_ESHAPE, "empty tile its..ite, jts..jte");
    const int nb = dt_partial_blocks();
    if (!h->d_dt) {
        rc = dalloc(h, (void **)&h->d_dt, sizeof(float) * ((size_t)d.nz + 4 * (size_t)nb));
        if (rc) return rc;
    }
    float *part = h->d_dt + d.nz;
    HIPCHK(h, hipMemcpyAsync(h->d_dt, dz_levels, sizeof(float) * d.nz, hipMemcpyHostToDevice,
                             h->stream));
    // reads u, v and w once (the tile's second read hits the caches)
    timed(h, K_DT, 4.0 * ((double)h->Nu + (double)h->Nv + (double)h->N), [&] {
        launch_dt_partials(h->stream, d, its, ite, jts, jte, g.dx,
                           h->dev_field[ICAR_ADV_FIELD_U], h->dev_field[ICAR_ADV_FIELD_V],
                           h->dev_field[ICAR_ADV_FIELD_W], h->d_dt, part);
    });
    HIPCHK(h, hipGetLastError());
    std::vector<float> hp(4 * (size_t)nb);
    HIPCHK(h, hipMemcpyAsync(hp.data(), part, sizeof(float) * hp.size(), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float mu = 0.0f, mv = 0.0f, mw = 0.0f, m3 = 0.0f;
    for (int b = 0; b < nb; b++) {
        mu = std::max(mu, hp[4 * b]);
        mv = std::max(mv, hp[4 * b + 1]);
        mw = std::max(mw, hp[4 * b + 2]);
        m3 = std::max(m3, hp[4 * b + 3]);
    }
    // time_step.f90:70-149, in FP32 as the reference computes it
    const float three_d_cfl = 0.577350269f;
    const float sqrt3 = sqrtf(3.0f) * 1.001f;
    float maxwind3d;
    if (cfl_strictness == 1) {
        maxwind3d = std::max(std::max(mu, mv), mw) * sqrt3;
    } else if (cfl_strictness == 5) {
        maxwind3d = mu + mv + mw;
    } else {
        maxwind3d = m3;
        if (cfl_strictness == 2)
            maxwind3d = std::max(std::max(std::max(mu, mv), mw), maxwind3d * three_d_cfl);
        else if (cfl_strictness == 4)
            maxwind3d = maxwind3d * sqrt3;
    }
    *dt = cfl / maxwind3d;
    if (*dt < 1e-1f)  // :158-166, where the reference stops
        return set_err(h, ICAR_ADV_ECOURANT,
                       "time step too small: dt = %g (Umax = %g, Vmax = %g, Wmax = %g)",
                       (double)*dt, (double)mu, (double)mv, (double)mw);
    return 0;
}

}  // extern "C"

namespace {

// init_relax_filters, src/objects/domain_obj.f90:3285-3390: the 2-D
// relaxation filter f[j*nx + i] for a domain with halo hs and the given
// boundary flags, statement by statement (later statements overwrite
// earlier ones; a domain touching boundaries in x and y takes the corner
// branches, of which only the first matching pair applies, so a one-image
// domain relaxes its north and west edges).
void relax_filter(int nx, int ny, int hs, bool west, bool east, bool north, bool south,
                  std::vector<float> &f)
{
    static const float rs[7] = {0.9f, 0.75f, 0.6f, 0.5f, 0.4f, 0.25f, 0.1f};
    static const float rs_r[7] = {0.1f, 0.25f, 0.4f, 0.5f, 0.6f, 0.75f, 0.9f};
    f.assign((size_t)nx * ny, 0.0f);
    auto at = [&](int i, int j) -> float & { return f[(size_t)j * nx + i]; };
    const int nr = std::min(7, std::min(nx - 1 - hs, ny - 1 - hs));  // :3302
    auto cols = [&](int i0, int i1, int j0, int j1) {  // value 1 on [i0,i1) x [j0,j1)
        for (int j = std::max(j0, 0); j < std::min(j1, ny); j++)
            for (int i = std::max(i0, 0); i < std::min(i1, nx); i++) at(i, j) = 1.0f;
    };
    auto ramp_x = [&](int i0, const float *r, int j0, int j1) {  // at(i0+t, j) = r[t]
        for (int j = std::max(j0, 0); j < std::min(j1, ny); j++)
            for (int t = 0; t < nr; t++)
                if (i0 + t >= 0 && i0 + t < nx) at(i0 + t, j) = r[t];
    };
    auto ramp_y = [&](int j0, const float *r, int i0, int i1) {  // at(i, j0+t) = r[t]
        for (int i = std::max(i0, 0); i < std::min(i1, nx); i++)
            for (int t = 0; t < nr; t++)
                if (j0 + t >= 0 && j0 + t < ny) at(i, j0 + t) = r[t];
    };
    auto square = [&](bool xlo, bool ylo) {  // the corner square, rs(min(a, b))
        for (int a = 1; a <= nr; a++)
            for (int b = 1; b <= nr; b++) {
                const int i = xlo ? hs + a - 1 : nx - hs - a;
                const int j = ylo ? hs + b - 1 : ny - hs - b;
                if (i >= 0 && i < nx && j >= 0 && j < ny) at(i, j) = rs[std::min(a, b) - 1];
            }
    };
    const bool corner = (west || east) && (north || south);
    if (!corner) {
        if (west) {
            cols(0, hs, 0, ny);
            ramp_x(hs, rs, 0, ny);
        } else if (east) {
            cols(nx - hs, nx, 0, ny);
            ramp_x(nx - hs - nr, rs_r, 0, ny);
        } else if (north) {
            cols(0, nx, ny - hs - 2, ny);  // rows jme-hs-1:jme (:3320)
            ramp_y(ny - hs - nr, rs_r, 0, nx);
        } else if (south) {
            cols(0, nx, 0, hs);
            ramp_y(hs, rs, 0, nx);
        }
    } else if (north && west) {
        cols(0, hs, 0, ny);
        cols(0, nx, ny - hs - 2, ny);
        ramp_x(hs, rs, 0, ny - hs);
        ramp_y(ny - hs - nr, rs_r, hs, nx);
        square(true, false);
    } else if (north && east) {
        cols(nx - hs, nx, 0, ny);
        cols(0, nx, ny - hs - 2, ny);
        ramp_x(nx - hs - nr, rs_r, 0, ny - hs);
        ramp_y(ny - hs - nr, rs_r, 0, nx - hs);
        square(false, false);
    } else if (south && west) {
        cols(0, hs, 0, ny);
        cols(0, nx, 0, hs);
        ramp_x(hs, rs, hs, ny);
        ramp_y(hs, rs, hs, nx);
        square(true, true);
    } else if (south && east) {
        cols(nx - hs, nx, 0, ny);
        cols(0, nx, 0, hs);
        ramp_x(nx - hs - nr, rs_r, hs, ny);
        ramp_y(hs, rs, 0, nx - hs);
        square(false, true);
    }
}

}  // namespace

extern "C" {

int icar_adv_set_relax_filter(void *handle, int west, int east, int north, int south, int halo)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (halo < 0 || halo >= std::min(h->d.nx, h->d.ny) - 1)
        return set_err(h, ICAR_ADV_ERANGE, "halo %d out of range", halo);
    std::vector<float> f;
    relax_filter(h->d.nx, h->d.ny, halo, west != 0, east != 0, north != 0, south != 0, f);
    if (!h->d_filter) {
        rc = dalloc(h, (void **)&h->d_filter, sizeof(float) * f.size());
        if (rc) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_filter, f.data(), sizeof(float) * f.size(),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_set_forcing(void *handle, int idx, const float *forcing_hi, const float *dqdt)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (idx < 0 || idx >= h->nspecies || !h->sp[idx].bound)
        return set_err(h, ICAR_ADV_ERANGE, "species %d is not bound", idx);
    if (!forcing_hi || !dqdt) return set_err(h, ICAR_ADV_EINVAL, "NULL forcing array");
    SpeciesSlot &s = h->sp[idx];
    if (s.storage != ICAR_ADV_STORE_F32)
        return set_err(h, ICAR_ADV_EINVAL, "lateral forcing needs an FP32-stored species");
    if (!s.fhi) {
        rc = dalloc(h, (void **)&s.fhi, sizeof(float) * h->N);
        if (!rc) rc = dalloc(h, (void **)&s.dqdt, sizeof(float) * h->N);
        if (rc) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(s.fhi, forcing_hi, sizeof(float) * h->N, hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(s.dqdt, dqdt, sizeof(float) * h->N, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_apply_forcing(void *handle, double dt)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!h->d_filter)
        return set_err(h, ICAR_ADV_EUNBOUND, "icar_adv_set_relax_filter has not been called");
    // the forcing kernel updates FP32 species: a forced species bound again
    // with 16-bit storage is refused before anything is launched
    for (int s = 0; s < h->nspecies; s++)
        if (h->sp[s].bound && h->sp[s].fhi && h->sp[s].storage != ICAR_ADV_STORE_F32)
            return set_err(h, ICAR_ADV_EINVAL,
                           "species %d has lateral forcing but 16-bit storage (forcing needs FP32)",
                           s);
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!sl.bound || !sl.fhi) continue;
        // reads the forcing, its tendency and the species, writes both fields
        timed(h, K_FORCING, 5.0 * 4 * cells(h), [&] {
            launch_forcing(h->stream, h->d, dt, h->d_filter, static_cast<float *>(species_dev(h, s)),
                           sl.fhi, sl.dqdt);
        });
        HIPCHK(h, hipGetLastError());
    }
    if (h->o.debug) HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int icar_adv_halo_bytes(void *handle, int side, unsigned long long *bytes)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (!bytes) return set_err(h, ICAR_ADV_EINVAL, "NULL output");
    Box snd, rcv;
    if (!side_boxes(h, side, &snd, &rcv))
        return set_err(h, ICAR_ADV_EINVAL, "side %d: unknown, or no halo on that side", side);
    HaloStrips hs{};
    *bytes = strip_table(h, snd, &hs, true);
    return 0;
}

int icar_adv_halo_pack(void *handle, int side, void *dev_buf, unsigned long long bytes)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    Box snd, rcv;
    if (!side_boxes(h, side, &snd, &rcv))
        return set_err(h, ICAR_ADV_EINVAL, "side %d: unknown, or no halo on that side", side);
    HaloStrips hs{};
    const size_t need = strip_table(h, snd, &hs, true);
    if (!dev_buf || bytes < need)
        return set_err(h, ICAR_ADV_ERANGE, "halo buffer of %llu bytes, side %d needs %zu", bytes,
                       side, need);
    hs.ni = snd.ni;
    hs.nj = snd.nj;
    hs.a_i0 = snd.i0;
    hs.a_j0 = snd.j0;
    hs.a_nx = h->d.nx;
    hs.buf = dev_buf;
    timed(h, K_HALO, 2.0 * need, [&] { launch_strip(h->stream, hs, h->d.nz, 0); });
    HIPCHK(h, hipGetLastError());
    return 0;
}

int icar_adv_halo_unpack(void *handle, int side, const void *dev_buf, unsigned long long bytes)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    Box snd, rcv;
    if (!side_boxes(h, side, &snd, &rcv))
        return set_err(h, ICAR_ADV_EINVAL, "side %d: unknown, or no halo on that side", side);
    HaloStrips hs{};
    const size_t need = strip_table(h, rcv, &hs, false);
    if (!dev_buf || bytes < need)
        return set_err(h, ICAR_ADV_ERANGE, "halo buffer of %llu bytes, side %d needs %zu", bytes,
                       side, need);
    hs.ni = rcv.ni;
    hs.nj = rcv.nj;
    hs.b_i0 = rcv.i0;
    hs.b_j0 = rcv.j0;
    hs.b_nx = h->d.nx;
    hs.buf = const_cast<void *>(dev_buf);
    timed(h, K_HALO, 2.0 * need, [&] { launch_strip(h->stream, hs, h->d.nz, 1); });
    HIPCHK(h, hipGetLastError());
    return 0;
}

int icar_adv_halo_exchange(void *from, int side, void *to)
{
    Handle *a, *b;
    int rc = check_handle(from, &a);
    if (rc) return rc;
    rc = check_handle(to, &b);
    if (rc) return rc;
    Box snd, rcv_a, snd_b, rcv;
    if (!side_boxes(a, side, &snd, &rcv_a) || !side_boxes(b, opposite(side), &snd_b, &rcv))
        return set_err(a, ICAR_ADV_EINVAL, "side %d: unknown, or no halo on one of the two sides",
                       side);
    HaloStrips hs{};
    strip_table(a, snd, &hs, true);
    HaloStrips hb{};
    strip_table(b, rcv, &hb, false);
    if (snd.ni != rcv.ni || snd.nj != rcv.nj || a->d.nz != b->d.nz || hs.n != hb.n)
        return set_err(a, ICAR_ADV_ESHAPE,
                       "neighbours do not match: strip %d x %d x %d (%d species) against %d x %d x "
                       "%d (%d species)",
                       snd.ni, a->d.nz, snd.nj, hs.n, rcv.ni, b->d.nz, rcv.nj, hb.n);
    // the strips carry stored values: both sides need each species in the
    // same format (FP32, BF16 or FP16; the rounding kind does not matter) and,
    // for 16-bit formats, with the same 2^e scale
    auto format = [](int st) {
        return st == ICAR_ADV_STORE_F32 ? 0
               : (st == ICAR_ADV_STORE_BF16 || st == ICAR_ADV_STORE_BF16_SR) ? 1
                                                                              : 2;
    };
    std::vector<const SpeciesSlot *> sa, sb;
    std::vector<int> ia, ib;  // species index of each strip
    for (int s = 0; s < a->nspecies; s++)
        if (a->sp[s].bound) sa.push_back(&a->sp[s]), ia.push_back(s);
    for (int s = 0; s < b->nspecies; s++)
        if (b->sp[s].bound) sb.push_back(&b->sp[s]), ib.push_back(s);
    for (int m = 0; m < hs.n; m++) {
        if (ia[m] != ib[m])
            return set_err(a, ICAR_ADV_ESHAPE,
                           "strip %d carries species %d on the sending side and species %d on "
                           "the receiving side: both sides need the same species bound",
                           m, ia[m], ib[m]);
        const int fa = format(sa[m]->storage), fb = format(sb[m]->storage);
        if (fa != fb || (fa != 0 && sa[m]->pow2 != sb[m]->pow2))
            return set_err(a, ICAR_ADV_ESHAPE,
                           "species %d is stored differently on the two sides (format %d, 2^%d "
                           "against format %d, 2^%d)",
                           m, fa, sa[m]->pow2, fb, sb[m]->pow2);
        hs.b[m] = hb.b[m];
    }
    hs.ni = snd.ni;
    hs.nj = snd.nj;
    hs.a_i0 = snd.i0;
    hs.a_j0 = snd.j0;
    hs.a_nx = a->d.nx;
    hs.b_i0 = rcv.i0;
    hs.b_j0 = rcv.j0;
    hs.b_nx = b->d.nx;
    // the copy runs on the receiver's stream after the sender's queued work;
    // the sender's next work waits for it
    hipEvent_t ea = pool_event(a), eb = pool_event(b);
    if (!ea || !eb) return set_err(a, ICAR_ADV_ENOMEM, "event creation failed");
    HIPCHK(a, hipEventRecord(ea, a->stream));
    HIPCHK(b, hipStreamWaitEvent(b->stream, ea, 0));
    timed(b, K_HALO, 2.0 * 4 * (double)snd.ni * snd.nj * a->d.nz * hs.n,
          [&] { launch_strip(b->stream, hs, b->d.nz, 2); });
    HIPCHK(b, hipGetLastError());
    HIPCHK(b, hipEventRecord(eb, b->stream));
    HIPCHK(a, hipStreamWaitEvent(a->stream, eb, 0));
    // a wait holds the event's state at the time it was enqueued, so the
    // events can be recorded again by later calls
    a->event_pool.push_back(ea);
    b->event_pool.push_back(eb);
    return 0;
}

int icar_adv_enforce_limits(void *handle)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!sl.bound) continue;
        const double es = sl.storage == ICAR_ADV_STORE_F32 ? 4.0 : 2.0;
        timed(h, K_CLAMP, 2.0 * es * cells(h),
              [&] { launch_clamp(h->stream, h->N, species_dev(h, s), sl.storage); });
    }
    HIPCHK(h, hipGetLastError());
    return 0;
}

int icar_adv_snapshot(void *handle, unsigned long long mask)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    rc = finish_snapshot(h);
    if (rc) return rc;
    // what goes out, in FP32 words
    std::vector<Handle::SnapItem> items;
    size_t total = 0;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        items.push_back({h->host_field[f], total, field_elems(h, f)});
        total += field_elems(h, f);
    }
    for (int s = 0; s < h->nspecies; s++) {
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !h->sp[s].bound) continue;
        items.push_back({h->sp[s].host, total, h->N});
        total += h->N;
    }
    if (items.empty()) return 0;
    if (total > h->snap_cap) {
        (void)hipFree(h->snap_dev);
        if (h->snap_host) (void)hipHostFree(h->snap_host);
        h->snap_dev = nullptr;
        h->snap_host = nullptr;
        h->snap_cap = 0;
        rc = dalloc(h, (void **)&h->snap_dev, total * sizeof(float));
        if (rc) return rc;
        // non-coherent: cacheable for the CPU copy into the host arrays
        if (hipHostMalloc((void **)&h->snap_host, total * sizeof(float),
                          hipHostMallocNonCoherent) != hipSuccess) {
            h->snap_host = nullptr;
            return set_err(h, ICAR_ADV_ENOMEM, "pinned host staging of %zu bytes failed",
                           total * sizeof(float));
        }
        h->snap_cap = total;
    }
    // each created on first use (a failed creation is retried by the next call)
    if (!h->copy_stream) HIPCHK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    if (!h->snap_ready) HIPCHK(h, hipEventCreateWithFlags(&h->snap_ready, hipEventDisableTiming));
    if (!h->snap_done) HIPCHK(h, hipEventCreateWithFlags(&h->snap_done, hipEventDisableTiming));
    // on the compute stream: the current state into the device staging
    size_t m = 0;
    for (int f = 0; f < ICAR_ADV_NFIELDS; f++) {
        if (!(mask & ICAR_ADV_MASK_FIELD(f)) || !h->host_field[f]) continue;
        const auto &it = items[m++];
        HIPCHK(h, hipMemcpyAsync(h->snap_dev + it.off, h->dev_field[f], it.n * sizeof(float),
                                 hipMemcpyDeviceToDevice, h->stream));
    }
    for (int s = 0; s < h->nspecies; s++) {
        SpeciesSlot &sl = h->sp[s];
        if (!(mask & ICAR_ADV_MASK_SPECIES(s)) || !sl.bound) continue;
        const auto &it = items[m++];
        timed(h, K_SNAPSHOT, 2.0 * 4 * cells(h), [&] {
            if (sl.storage == ICAR_ADV_STORE_F32)
                (void)hipMemcpyAsync(h->snap_dev + it.off, species_dev(h, s), it.n * sizeof(float),
                                     hipMemcpyDeviceToDevice, h->stream);
            else
                launch_unpack(h->stream, h->N, species_dev(h, s), h->snap_dev + it.off, sl.storage,
                              sl.pow2);
        });
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->snap_ready, h->stream));
    // on the copy stream, overlapping whatever the compute stream does next
    HIPCHK(h, hipStreamWaitEvent(h->copy_stream, h->snap_ready, 0));
    HIPCHK(h, hipMemcpyAsync(h->snap_host, h->snap_dev, total * sizeof(float),
                             hipMemcpyDeviceToHost, h->copy_stream));
    HIPCHK(h, hipEventRecord(h->snap_done, h->copy_stream));
    h->snap_items = items;
    h->snap_pending = true;
    h->snap_mask = mask;
    return 0;
}

int icar_adv_snapshot_wait(void *handle)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    return finish_snapshot(h);
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
    return apply_bc_on(h, h, h->stream, bc_x, bc_y);
}

int icar_adv_run(void *handle, float dt, int nsteps, int bc_x, int bc_y)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    if (nsteps < 0) return set_err(h, ICAR_ADV_EINVAL, "nsteps < 0");
    const bool fused = h->fast && h->o.kernel_variant == 1 && !h->o.rk3 && h->nspecies > 0 &&
                       h->o.mpdata_order >= 1;
    if (bc_x < 0 || bc_x > 2 || (bc_y != 0 && bc_y != 2))
        return set_err(h, ICAR_ADV_EINVAL, "bc_x must be 0/1/2 and bc_y 0/2");
    auto steps = [&]() -> int {
        for (int s = 0; s < nsteps; s++) {
            int r = do_advect(h, dt);
            if (r) return r;
            if (fused && h->batch) {  // every species' ring conditions in two launches
                timed(h, K_BC, 0.0, [&] { r = fast_batch_bc(h->batch, h->stream, bc_x, bc_y); });
                if (r) return set_err(h, r, "ring conditions: %s", hipGetErrorString((hipError_t)r));
            } else {
                r = icar_adv_apply_bc(handle, bc_x, bc_y);
                if (r) return r;
            }
        }
        return 0;
    };
    // graph capture needs every launch to be asynchronous: not with the
    // profiling events or the debug check's host copies
    if (!h->use_graph || h->prof || h->o.debug || nsteps == 0) return steps();
    const int p0 = parity(h);
    if (h->gexec && h->g_dt == dt && h->g_steps == nsteps && h->g_bcx == bc_x &&
        h->g_bcy == bc_y && h->g_par0 == p0) {
        set_parity(h, h->g_par1);
        HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
        return 0;
    }
    if (h->gexec) {
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    rc = require_bound(h);
    if (rc) return rc;
    if (fused) {  // table uploads happen before the capture
        rc = ensure_batch(h);
        if (rc) return rc;
        std::string e;
        rc = fast_batch_ready(h->batch, h->stream, &dt, &e);  // autotuning, if due, runs here
        if (rc) return set_err(h, rc, "%s", e.c_str());
    }
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = steps();
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_parity(h, p0);
        if (rc) return rc;
        return set_err(h, (int)ec, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    }
    ec = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ec != hipSuccess) {
        h->gexec = nullptr;
        set_parity(h, p0);
        return set_err(h, (int)ec, "hipGraphInstantiate failed: %s", hipGetErrorString(ec));
    }
    h->g_dt = dt;
    h->g_steps = nsteps;
    h->g_bcx = bc_x;
    h->g_bcy = bc_y;
    h->g_par0 = p0;
    h->g_par1 = parity(h);  // advanced while recording
    HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    return 0;
}

int icar_adv_set_graph(void *handle, int enable)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    h->use_graph = enable != 0;
    return 0;
}

int icar_adv_launch_config(void *handle, int *corr_cfg, int *ksplit)
{
    Handle *h;
    int rc = check_handle(handle, &h);
    if (rc) return rc;
    fast_batch_config(h->batch, corr_cfg, ksplit);
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

// ---------------------------------------------------------------------------
// Ensembles: independent domains advanced by one launch per pass
// ---------------------------------------------------------------------------
namespace {
struct Ensemble {
    std::vector<Handle *> hs;
    std::vector<unsigned> gens;  // each member's bind_gen when the batch was built
    FastBatch *batch = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    bool prof = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double acc_ms[K_COUNT] = {};
    long long acc_n[K_COUNT] = {};
    double acc_bytes[K_COUNT] = {};
    hipGraphExec_t gexec = nullptr;
    std::vector<float> g_dt;
    int g_steps = -1, g_bcx = -1, g_bcy = -1, g_par0 = -1, g_par1 = -1;
};

int ens_err(Ensemble *e, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e)
        e->err = buf;
    else
        g_create_error = buf;
    return code;
}

void ens_free(Ensemble *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drain_profile(e);
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    fast_batch_destroy(e->batch);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int ens_steps(Ensemble *e, const float *dt, int nsteps, int bc_x, int bc_y)
{
    if (bc_x < 0 || bc_x > 2 || (bc_y != 0 && bc_y != 2))
        return ens_err(e, ICAR_ADV_EINVAL, "bc_x must be 0/1/2 and bc_y 0/2");
    for (int st = 0; st < nsteps; st++) {
        std::string m;
        int rc = fast_batch_advect(e->batch, e->stream, dt, timed_fn(e), &m);
        if (rc) return ens_err(e, rc, "%s", m.c_str());
        for (size_t k = 0; k < e->hs.size(); k++) {
            e->hs[k]->last_dt = dt[k];
            e->hs[k]->have_dt = true;
        }
        // the ring conditions of every member's species: two launches
        timed(e, K_BC, 0.0, [&] { rc = fast_batch_bc(e->batch, e->stream, bc_x, bc_y); });
        if (rc) return ens_err(e, rc, "ring conditions: %s", hipGetErrorString((hipError_t)rc));
    }
    hipError_t x = hipGetLastError();
    if (x != hipSuccess) return ens_err(e, (int)x, "launch failed: %s", hipGetErrorString(x));
    return 0;
}

void ens_set_parity(Ensemble *e, int p)
{
    for (Handle *h : e->hs) fast_set_parity(h->fast, p);
}
}  // namespace

extern "C" {

int icar_adv_ensemble_create(int n, void *const *handles, void **ens)
{
    if (!ens || !handles || n < 1) return ens_err(nullptr, ICAR_ADV_EINVAL, "bad ensemble arguments");
    *ens = nullptr;
    std::vector<FastDomain> doms;
    Ensemble *e = new Ensemble();
    for (int k = 0; k < n; k++) {
        Handle *h = static_cast<Handle *>(handles[k]);
        if (!h) {
            ens_free(e);
            return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL handle %d", k);
        }
        // one launch updates every member: a domain listed twice would be
        // advanced twice over the same buffers at once
        for (int q = 0; q < k; q++)
            if (handles[q] == handles[k]) {
                ens_free(e);
                return ens_err(nullptr, ICAR_ADV_EINVAL, "handle %d repeats handle %d", k, q);
            }
        if (h->is_std || h->o.kernel_variant != 1 || h->o.rk3 || h->o.mpdata_order < 1 ||
            h->nspecies == 0 || require_bound(h) != 0) {
            ens_free(e);
            return ens_err(nullptr, ICAR_ADV_EINVAL,
                           "handle %d: ensembles need MPDATA, kernel_variant 1, RK3 off, "
                           "mpdata_order >= 1, bound fields and species",
                           k);
        }
        if (h->o.mpdata_order != static_cast<Handle *>(handles[0])->o.mpdata_order ||
            h->o.mpdata_fct != static_cast<Handle *>(handles[0])->o.mpdata_fct) {
            ens_free(e);
            return ens_err(nullptr, ICAR_ADV_EINVAL,
                           "all domains of an ensemble need the same mpdata_order and MPDATA_FCT");
        }
        (void)hipStreamSynchronize(h->stream);
        e->hs.push_back(h);
        e->gens.push_back(h->bind_gen);
        doms.push_back({h->fast, domain_fields(h), h->sf, h->g.dx, h->o.advect_density,
                        h->compensated});
    }
    std::string m;
    e->batch = fast_batch_create(doms.data(), n, e->hs[0]->o.mpdata_order, e->hs[0]->o.mpdata_fct, &m);
    if (!e->batch) {
        ens_free(e);
        return ens_err(nullptr, ICAR_ADV_EINVAL, "%s", m.c_str());
    }
    hipError_t x = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (x != hipSuccess) {
        e->stream = nullptr;
        ens_free(e);
        return ens_err(nullptr, (int)x, "hipStreamCreate failed: %s", hipGetErrorString(x));
    }
    *ens = e;
    return 0;
}

int icar_adv_ensemble_destroy(void *ens)
{
    ens_free(static_cast<Ensemble *>(ens));
    return 0;
}

int icar_adv_ensemble_run(void *ens, const float *dt, int nsteps, int bc_x, int bc_y, int use_graph)
{
    Ensemble *e = static_cast<Ensemble *>(ens);
    if (!e) return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL ensemble");
    if (!dt || nsteps < 0) return ens_err(e, ICAR_ADV_EINVAL, "bad dt or nsteps");
    const int n = (int)e->hs.size();
    for (int k = 0; k < n; k++)
        if (e->hs[k]->bind_gen != e->gens[k])
            return ens_err(e, ICAR_ADV_EINVAL,
                           "member %d bound a species after the ensemble was created: its "
                           "buffers may have moved (create the ensemble again)",
                           k);
    for (Handle *h : e->hs) (void)hipStreamSynchronize(h->stream);
    if (!use_graph || e->prof || nsteps == 0) return ens_steps(e, dt, nsteps, bc_x, bc_y);
    const int p0 = fast_parity(e->hs[0]->fast);
    bool hit = e->gexec && e->g_steps == nsteps && e->g_bcx == bc_x && e->g_bcy == bc_y &&
               e->g_par0 == p0 && (int)e->g_dt.size() == n;
    for (int k = 0; hit && k < n; k++) hit = e->g_dt[k] == dt[k];
    if (hit) {
        ens_set_parity(e, e->g_par1);
        hipError_t x = hipGraphLaunch(e->gexec, e->stream);
        return x == hipSuccess ? 0 : ens_err(e, (int)x, "hipGraphLaunch: %s", hipGetErrorString(x));
    }
    if (e->gexec) {
        (void)hipGraphExecDestroy(e->gexec);
        e->gexec = nullptr;
    }
    std::string m;
    int rc = fast_batch_ready(e->batch, e->stream, dt, &m);
    if (rc) return ens_err(e, rc, "%s", m.c_str());
    hipError_t x = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if (x != hipSuccess) return ens_err(e, (int)x, "hipStreamBeginCapture: %s", hipGetErrorString(x));
    rc = ens_steps(e, dt, nsteps, bc_x, bc_y);
    hipGraph_t g = nullptr;
    x = hipStreamEndCapture(e->stream, &g);
    if (rc || x != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        ens_set_parity(e, p0);
        return rc ? rc : ens_err(e, (int)x, "hipStreamEndCapture: %s", hipGetErrorString(x));
    }
    x = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (x != hipSuccess) {
        e->gexec = nullptr;
        ens_set_parity(e, p0);
        return ens_err(e, (int)x, "hipGraphInstantiate: %s", hipGetErrorString(x));
    }
    e->g_dt.assign(dt, dt + n);
    e->g_steps = nsteps;
    e->g_bcx = bc_x;
    e->g_bcy = bc_y;
    e->g_par0 = p0;
    e->g_par1 = fast_parity(e->hs[0]->fast);
    x = hipGraphLaunch(e->gexec, e->stream);
    return x == hipSuccess ? 0 : ens_err(e, (int)x, "hipGraphLaunch: %s", hipGetErrorString(x));
}

int icar_adv_ensemble_sync(void *ens)
{
    Ensemble *e = static_cast<Ensemble *>(ens);
    if (!e) return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL ensemble");
    hipError_t x = hipStreamSynchronize(e->stream);
    return x == hipSuccess ? 0 : ens_err(e, (int)x, "hipStreamSynchronize: %s", hipGetErrorString(x));
}

const char *icar_adv_ensemble_last_error(void *ens)
{
    if (!ens) return g_create_error.c_str();
    return static_cast<Ensemble *>(ens)->err.c_str();
}

int icar_adv_ensemble_profile(void *ens, int enable)
{
    Ensemble *e = static_cast<Ensemble *>(ens);
    if (!e) return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL ensemble");
    (void)hipStreamSynchronize(e->stream);
    drain_profile(e);
    if (enable)
        for (int k = 0; k < K_COUNT; k++) e->acc_ms[k] = e->acc_n[k] = e->acc_bytes[k] = 0;
    e->prof = enable != 0;
    return 0;
}

int icar_adv_ensemble_launch_config(void *ens, int *corr_cfg, int *ksplit)
{
    Ensemble *e = static_cast<Ensemble *>(ens);
    if (!e) return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL ensemble");
    fast_batch_config(e->batch, corr_cfg, ksplit);
    return 0;
}

int icar_adv_ensemble_kernel_stats(void *ens, int id, double *total_ms, long long *launches,
                                   double *alg_bytes)
{
    Ensemble *e = static_cast<Ensemble *>(ens);
    if (!e) return ens_err(nullptr, ICAR_ADV_EINVAL, "NULL ensemble");
    if (id < 0 || id >= K_COUNT) return ens_err(e, ICAR_ADV_EINVAL, "unknown kernel id %d", id);
    (void)hipStreamSynchronize(e->stream);
    drain_profile(e);
    if (total_ms) *total_ms = e->acc_ms[id];
    if (launches) *launches = e->acc_n[id];
    if (alg_bytes) *alg_bytes = e->acc_bytes[id];
    return 0;
}

}  // extern "C"

namespace icaradv {
int set_call_error(int code, const char *msg)
{
    g_create_error = msg;
    return code;
}

const char *kernel_name(int id)
{
    static const char *names[K_COUNT] = {"courant",     "upwind",     "pseudo_velocity",
                                         "fct",         "ring_bc",    "reduce",
                                         "copy",        "fused_first_pass", "fused_corrective_pass",
                                         "convert16",   "face_coefficients", "wind_balance",
                                         "dt_bound",    "forcing",    "std_wind",
                                         "std_stage",   "std_limiter_scale", "std_limiter_final",
                                         "std_ring",    "halo_strip", "enforce_limits",
                                         "snapshot"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "";
}
}  // namespace icaradv
