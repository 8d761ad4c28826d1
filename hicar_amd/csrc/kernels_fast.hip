// Fused species-batched path (kernel_variant 1): not yet built; the entry
// point reports a clear error.  Reduced-precision storage helpers live here.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <string>

#include "../../include/icar_adv.h"
#include "fast.hpp"

namespace icaradv {

struct FastState {
    Dims d;
    std::string err;
};

FastState *fast_create(Dims d)
{
    FastState *st = new FastState();
    st->d = d;
    return st;
}

void fast_destroy(FastState *st) { delete st; }

int fast_set_species(FastState *, const FastSpecies *, int) { return 0; }

const char *fast_error(FastState *st) { return st ? st->err.c_str() : "no fused state"; }

int fast_advect(FastState *st, hipStream_t, Dims, const StepFields &, const DomainFields &, int, int,
                const TimedFn &)
{
    st->err = "kernel_variant 1 (fused path) is not available in this build";
    return ICAR_ADV_EINVAL;
}

namespace {

__device__ __forceinline__ float load16(const uint16_t *p, size_t i, int storage, float scale)
{
    const uint16_t b = p[i];
    float v;
    if (storage == ICAR_ADV_STORE_BF16)
        v = __uint_as_float((uint32_t)b << 16);
    else
        v = __half2float(__ushort_as_half(b));
    return v * scale;
}

__device__ __forceinline__ uint16_t store16(float v, int storage, float inv_scale)
{
    const float x = v * inv_scale;
    if (storage == ICAR_ADV_STORE_BF16) {
        // round to nearest even
        uint32_t u = __float_as_uint(x);
        if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
    return __half_as_ushort(__float2half_rn(x));
}

__global__ void k_pack(size_t n, const float *__restrict__ src, uint16_t *__restrict__ dst,
                       int storage, float inv_scale)
{
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        dst[p] = store16(src[p], storage, inv_scale);
}

__global__ void k_unpack(size_t n, const uint16_t *__restrict__ src, float *__restrict__ dst,
                         int storage, float scale)
{
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x)
        dst[p] = load16(src, p, storage, scale);
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

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (unsigned)(b < 1 ? 1 : b);
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
