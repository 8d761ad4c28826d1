// Fused, species-batched MPDATA path (kernel_variant 1) and the reduced
// precision species storage it supports.
#pragma once
#include <functional>

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
int fast_set_species(FastState *st, const FastSpecies *sp, int n);
const char *fast_error(FastState *st);
// Device buffer holding the current state of species s (in bind order).
void *fast_species_ptr(FastState *st, int s);

using TimedFn = std::function<void(int id, double bytes, const std::function<void()> &)>;
int fast_advect(FastState *st, hipStream_t s, Dims d, const StepFields &sf, const DomainFields &f,
                int order, int fct, const TimedFn &timed);

// FP32 <-> scaled BF16/FP16 conversion and ring conditions for 16-bit storage.
void launch_pack(hipStream_t s, size_t n, const float *src, void *dst, int storage, int pow2);
void launch_unpack(hipStream_t s, size_t n, const void *src, float *dst, int storage, int pow2);
void launch_bc16(hipStream_t s, Dims d, void *q, int bc_x, int bc_y);

}  // namespace icaradv
