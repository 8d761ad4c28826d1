// The FP64 mass budget of one species without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// It is icar_adv_mass_budget's computation for one FP32 species
// (sum(q * jacobian * density * advection_dz) over the cells advect updates):
// the same k_mass kernel writes one partial sum per block, then one thread
// adds the partials in block order, as the handle's entry point does on the
// host, so the two give the same bits.  The partials live in a fixed array
// of this library on each device instead of the handle's buffer, the inputs
// and the one-double output are the caller's device memory, and the call
// only enqueues two kernels on the caller's stream: no allocation, no memset
// and no synchronisation, so it can be captured in a graph.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/icar_adv.h"
#include "common.hpp"

using namespace icaradv;

namespace {

// One partial sum per block of k_mass, which launches at most
// ICAR_ADV_NOALLOC_PARTIALS blocks; each GPU holds its own copy until the
// process ends.
__device__ __align__(256) double scratch[ICAR_ADV_NOALLOC_PARTIALS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};  // scratch's device address, per device

// The device address of `scratch` on the current device, looked up once per
// device (host code never passes the array's name itself to a kernel).
int scratch_of_current_device(double **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_call_error(ICAR_ADV_ERANGE, "device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            set_call_error((int)e, "hipGetSymbolAddress of the mass-budget scratch failed");
            return (int)e;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

// partial[0] + partial[1] + ... in block order (as the host loop of
// icar_adv_mass_budget adds them)
__global__ void k_sum_partials(const double *__restrict__ partial, int n, double *out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < n; b++) s += partial[b];
    *out = s;
}

}  // namespace

extern "C" int icar_adv_mass_budget_noalloc(int nx, int ny, int nz, const float *q,
                                            const float *jacobian, const float *density,
                                            const float *advection_dz, double *out,
                                            void *stream)
{
    if (nx < 3 || ny < 3 || nz < 3)
        return set_call_error(ICAR_ADV_EINVAL, "mass_budget_noalloc: nx, ny and nz must be >= 3");
    if (!q || !jacobian || !density || !advection_dz || !out)
        return set_call_error(ICAR_ADV_EINVAL, "mass_budget_noalloc: NULL array");
    const Dims d{nx, ny, nz};
    const int nb = mass_partials_blocks(d);
    if (nb > ICAR_ADV_NOALLOC_PARTIALS)
        return set_call_error(ICAR_ADV_ERANGE,
                              "mass_budget_noalloc: more partial sums than the fixed scratch holds");
    double *part = nullptr;
    const int rc = scratch_of_current_device(&part);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    launch_mass(s, d, q, jacobian, density, advection_dz, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, part, nb, out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        set_call_error((int)e, hipGetErrorString(e));
        return (int)e;
    }
    return 0;
}
