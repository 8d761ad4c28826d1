// Correctly rounded FP32 division in three operations once a reciprocal is
// known, for operands in a range where no step over- or underflows.
//
// With y the reciprocal of d and q0 = RN(x*y), the remainder x - q0*d is
// exactly representable and one FMA gives it; q = RN(q0 + (x - q0*d)*y) is
// then the correctly rounded quotient RN(x/d) when y is within half an ulp
// of 1/d and q0 within one ulp of x/d (Markstein, "IA-64 and Elementary
// Functions", 2000, theorem 8.10; Cornea, Harrison and Tang, "Scientific
// Computing on Itanium-based Systems", 2002, ch. 8).  The reciprocal here is
// the hardware estimate (v_rcp_f32, 1 ulp) refined by one Newton step.  The
// library does not rely on the theorem's premises holding for that y:
// icar_adv_verify_division compares the quotient with the IEEE division for
// every pair of significands on the device, and the kernels take this path
// only inside the range where every operation scales exactly with the
// operands' exponents (fdiv_in_range), so the significand pairs cover all of
// it.
//
// The remainder is formed negated, rn = q0*d - x, and the quotient as
// q0 - rn*y: for d > 0 a zero numerator then keeps its sign (+0 / d = +0,
// -0 / d = -0), as the IEEE division does.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace icaradv {

// 1/d: the hardware estimate and one Newton step, e = 1 - d*y0, y = y0 + e*y0
__device__ __forceinline__ float recip_nr(float d)
{
    const float y0 = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}

// RN(x/d) from y = recip_nr(d), for operands fdiv_in_range accepts
__device__ __forceinline__ float div_y(float x, float d, float y)
{
    const float q0 = x * y;
    const float rn = __builtin_fmaf(q0, d, -x);
    return __builtin_fmaf(-rn, y, q0);
}

// The range the fast quotient is taken in: d in [2^-60, 2^40] (positive,
// down to the 1e-15 the FCT limiter adds to a zero inflow), x zero or |x|
// in [2^-78, 2^64].  Then 1/d, x*y and the quotient are normal (|x/d| in
// [2^-118, 2^124]), and the remainder is zero or at least 2^-125 (x is a
// multiple of 2^(ex-23), q0*d of 2^(ex-48)), so it is exact and normal:
// every step equals, scaled by a power of two, the same step on the
// operands' significands.  The device check covers the reciprocal's scaling
// over d's exponents FDIV_E_LO..FDIV_E_HI.
constexpr int FDIV_E_LO = -60, FDIV_E_HI = 40;
constexpr float FDIV_D_LO = 0x1p-60f, FDIV_D_HI = 0x1p40f;
constexpr float FDIV_X_LO = 0x1p-78f, FDIV_X_HI = 0x1p64f;

// Guard accumulators of a group of fast quotients, checked once for the
// group: the smallest and largest divisor, the largest |numerator| and the
// smallest non-zero |numerator| (as 2*bits - 1 unsigned, so that +-0 maps to
// the largest value and never counts as small).
struct DivGuard {
    float dlo = FDIV_D_HI, dhi = FDIV_D_LO, xhi = 0.0f;
    uint32_t xlo = 0xffffffffu;
};

__device__ __forceinline__ uint32_t nz_key(float x)
{
    return (__float_as_uint(x) << 1) - 1u;  // +-0 -> 0xffffffff
}

// a division's operands into the guard
__device__ __forceinline__ void guard_div(DivGuard &g, float x, float d)
{
    g.dlo = fminf(g.dlo, d);
    g.dhi = fmaxf(g.dhi, d);
    g.xhi = fmaxf(g.xhi, fabsf(x));
    g.xlo = min(g.xlo, nz_key(x));
}

// a numerator only (its divisor is known to be in range)
__device__ __forceinline__ void guard_num(DivGuard &g, float x)
{
    g.xhi = fmaxf(g.xhi, fabsf(x));
    g.xlo = min(g.xlo, nz_key(x));
}

// a divisor only
__device__ __forceinline__ void guard_den(DivGuard &g, float d)
{
    g.dlo = fminf(g.dlo, d);
    g.dhi = fmaxf(g.dhi, d);
}

// true when every guarded quotient of this lane was in range (NaN operands
// pass: the quotient is NaN either way)
__device__ __forceinline__ bool guard_ok(const DivGuard &g)
{
    return g.dlo >= FDIV_D_LO && g.dhi <= FDIV_D_HI && g.xhi <= FDIV_X_HI &&
           g.xlo >= ((__float_as_uint(FDIV_X_LO) << 1) - 1u);
}

}  // namespace icaradv
