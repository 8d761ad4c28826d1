// Fused, species-batched MPDATA path (kernel_variant 1) and the reduced
// precision species storage it supports.
#pragma once
#include <functional>
#include <string>

#include "common.hpp"

namespace icaradv {

struct FastSpecies {
    void *dev;    // float* (FP32) or uint16_t* (BF16 / FP16, value * 2^-pow2)
    int storage;  // ICAR_ADV_STORE_*
    int pow2;
};

struct FastState;

FastState *fast_create(Dims d);
void fast_destroy(FastState *st);
// Replaces the species list.  Species that keep their home buffer keep
// their current device values (copied home on `stream` when the current
// state is in the second buffer); the second buffers are allocated anew.
int fast_set_species(FastState *st, const FastSpecies *sp, int n, hipStream_t stream);
const char *fast_error(FastState *st);
// Device buffer holding the current state of species s (in bind order).
void *fast_species_ptr(FastState *st, int s);

// Current home-buffer parity of a domain (alternates each step for
// mpdata_order <= 2).
int fast_parity(FastState *st);
void fast_set_parity(FastState *st, int parity);

// A launch group: one domain (a handle) or an ensemble of domains advanced
// by one launch per pass.
struct FastDomain {
    FastState *st;
    DomainFields f;
    StepFields sf;
    float dx;
    int advect_density;
    int compensated;  // the flux divergence with compensated sums (icar_adv_set_compensation)
};
struct FastBatch;
FastBatch *fast_batch_create(const FastDomain *doms, int n, int order, int fct, std::string *err);
void fast_batch_destroy(FastBatch *b);
// Uploads the domain tables for these time steps (one per domain) if they
// changed; fast_batch_advect calls it, graph capture calls it beforehand.
int fast_batch_prepare(FastBatch *b, const float *dt, std::string *err);

// Chooses the corrective kernel's launch configuration and column chunking
// for this batch by timing the candidates once (the result is cached by
// batch shape for the process); a no-op once chosen or while the stream is
// being captured.  fast_batch_ready = prepare + tune; call it before a graph
// capture.
int fast_batch_tune(FastBatch *b, hipStream_t s, std::string *err);
int fast_batch_ready(FastBatch *b, hipStream_t s, const float *dt, std::string *err);
// The chosen configuration (corr_cfg 0..3) and the column chunks of the
// first domain.
void fast_batch_config(const FastBatch *b, int *corr_cfg, int *ksplit);

using TimedFn = std::function<void(int id, double bytes, const std::function<void()> &)>;
int fast_batch_advect(FastBatch *b, hipStream_t s, const float *dt, const TimedFn &timed,
                      std::string *err);
// The ring conditions (bc_x 0/1/2, bc_y 0/2, as icar_adv_apply_bc) on the
// current buffer of every species of every domain of the batch: two launches.
int fast_batch_bc(const FastBatch *b, hipStream_t s, int bc_x, int bc_y);

// FP32 <-> scaled BF16/FP16 conversion and ring conditions for 16-bit storage.
void launch_pack(hipStream_t s, size_t n, const float *src, void *dst, int storage, int pow2);
void launch_unpack(hipStream_t s, size_t n, const void *src, float *dst, int storage, int pow2);
void launch_bc16(hipStream_t s, Dims d, void *q, int bc_x, int bc_y);

}  // namespace icaradv
