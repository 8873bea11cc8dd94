// x / d for the integer divisors d = 1..6 of the red-black update, as one
// multiply by y = RN(1/d) and one FMA residual step (Markstein's correction
// of a reciprocal quotient).  For every finite x with |x| >= 2^-123 the result
// is the correctly rounded quotient x / d: then x / d >= 2^-123 / 6 > 2^-126
// is a normal number and the residual x - q0 d is exact.  (Checked against
// the IEEE division over every FP32 x of 2^-127 <= |x| < 2^-120 and a random
// sample of all others, for each d, by tests/test_host.py.)  Tiny nonzero,
// infinite and NaN x take the IEEE division; copysign keeps -0 for x = -0.
// The header compiles as host C++ too, so the same function is checked on the
// CPU.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define FLUB_HD __host__ __device__ __forceinline__
#else
#define FLUB_HD static inline
#endif

FLUB_HD float div_small(float x, float df, float y)
{
    const float q0 = x * y;
    float q = copysignf(fmaf(fmaf(-q0, df, x), y, q0), x);
    const float ax = fabsf(x);
    if ((ax < 0x1p-123f && ax != 0.0f) || !(ax <= 0x1.fffffep127f)) q = x / df;
    return q;
}

#if defined(__HIPCC__)
// The same quotient in device code with the IEEE division behind a
// wavefront-uniform branch: written as above, the compiler computes both
// quotients for every lane and selects, and the IEEE division is most of
// the instructions of a cell update.
__device__ __forceinline__ float div_small_wave(float x, float df, float y)
{
    const float q0 = x * y;
    float q = copysignf(fmaf(fmaf(-q0, df, x), y, q0), x);
    const float ax = fabsf(x);
    const bool slow = (ax < 0x1p-123f && ax != 0.0f) || !(ax <= 0x1.fffffep127f);
    if (__builtin_amdgcn_ballot_w64(slow) != 0ull) {
        float xs = x;
        asm volatile("" : "+v"(xs));   // keeps the division inside the branch
        if (slow) q = xs / df;
    }
    return q;
}
#endif
