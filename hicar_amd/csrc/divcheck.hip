// Device check of the fast quotient of fdiv.hpp against the IEEE division,
// over every pair of significands (icar_adv_verify_division).  The kernels
// take the fast quotient only where each of its operations scales exactly
// with the operands' exponents (fdiv_in_range), so equality for x, d in
// [1, 2) together with the reciprocal's exact scaling by d's exponent covers
// every quotient the kernels take that way.  Compiled without contraction;
// `x / d` here is the IEEE division.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/icar_adv.h"
#include "common.hpp"
#include "fdiv.hpp"

using namespace icaradv;

namespace {

constexpr unsigned SIG = 1u << 23;   // significands per binade
constexpr unsigned XSPLIT = 16;      // threads sharing one divisor
constexpr unsigned CHECK_THREADS = 256;

__device__ __forceinline__ void report(unsigned long long *bad, unsigned long long *first,
                                       unsigned long long key, unsigned long long n)
{
    atomicAdd(bad, n);
    atomicCAS(first, ~0ull, key);
}

// thread t: divisor significand d_first + t / XSPLIT, numerators
// (t % XSPLIT) + XSPLIT * k; both in [1, 2).  Also, for its divisor, the
// reciprocal at every exponent e in [FDIV_E_LO, FDIV_E_HI] against the one
// at e = 0 scaled by 2^-e, and the quotient of -x (once per thread).
__global__ void __launch_bounds__(CHECK_THREADS)
    k_divcheck(unsigned d_first, unsigned d_count, unsigned long long *bad,
               unsigned long long *first)
{
    const unsigned t = blockIdx.x * CHECK_THREADS + threadIdx.x;
    if (t >= d_count * XSPLIT) return;
    const unsigned md = d_first + t / XSPLIT;
    const unsigned part = t % XSPLIT;
    const float d = __uint_as_float(0x3f800000u | md);
    const float y = recip_nr(d);
    unsigned long long nbad = 0, key = ~0ull;
    for (unsigned mx = part; mx < SIG; mx += XSPLIT) {
        const float x = __uint_as_float(0x3f800000u | mx);
        const float fast = div_y(x, d, y);
        const float ieee = x / d;
        if (__float_as_uint(fast) != __float_as_uint(ieee)) {
            nbad++;
            if (key == ~0ull) key = ((unsigned long long)md << 32) | mx;
        }
    }
    if (part == 0) {
        for (int e = FDIV_E_LO; e <= FDIV_E_HI; e++) {
            const float ds = ldexpf(d, e);
            if (__float_as_uint(recip_nr(ds)) != __float_as_uint(ldexpf(y, -e))) {
                nbad++;
                if (key == ~0ull) key = ((unsigned long long)md << 32) | 0x80000000ull | (e + 64);
            }
        }
        const float x = __uint_as_float(0xbf800000u | (md ^ 0x2aaaaau));  // a negative numerator
        if (__float_as_uint(div_y(x, d, y)) != __float_as_uint(x / d)) {
            nbad++;
            if (key == ~0ull) key = ((unsigned long long)md << 32) | 0x40000000ull;
        }
    }
    if (nbad) report(bad, first, key, nbad);
}

}  // namespace

extern "C" int icar_adv_verify_division(unsigned d_first, unsigned d_count,
                                        unsigned long long *mismatches, unsigned long long *first)
{
    if (!mismatches || !first || d_count == 0 || d_first >= SIG || d_count > SIG - d_first)
        return set_call_error(ICAR_ADV_EINVAL, "verify_division: bad range or NULL output");
    unsigned long long *dev = nullptr;
    hipError_t e = hipMalloc(&dev, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return set_call_error(ICAR_ADV_ENOMEM, "verify_division: hipMalloc failed");
    const unsigned long long init[2] = {0ull, ~0ull};
    e = hipMemcpy(dev, init, sizeof init, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned threads = d_count * XSPLIT;
        hipLaunchKernelGGL(k_divcheck, dim3((threads + CHECK_THREADS - 1) / CHECK_THREADS),
                           dim3(CHECK_THREADS), 0, nullptr, d_first, d_count, dev, dev + 1);
        e = hipGetLastError();
    }
    unsigned long long out[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(out, dev, sizeof out, hipMemcpyDeviceToHost);
    (void)hipFree(dev);
    if (e != hipSuccess) {
        set_call_error((int)e, hipGetErrorString(e));
        return (int)e;
    }
    *mismatches = out[0];
    *first = out[1];
    return 0;
}
