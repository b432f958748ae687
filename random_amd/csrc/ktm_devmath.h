/*
 * ktm_devmath.h -- correctly rounded division and square root used where a
 * kernel must reproduce the rounding of the reference's C arithmetic.
 */
#ifndef KTM_DEVMATH_H
#define KTM_DEVMATH_H

#include <hip/hip_runtime.h>

/* RN(a/b) from rb = RN(1/b): q = RN(a rb) is within one ulp of a/b, the
 * residual a - b q is exact under FMA, and one correction step rounds
 * correctly (Markstein).  Checked against true division on 5e8 pairs,
 * exhaustively over [2^-10, 32) for b = 0.004 (the DSR / dt division). */
__device__ __forceinline__ float div_cr(float a, float b, float rb)
{
    const float q = a * rb;
    const float r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, rb, q);
}

/* RN(a/b) in double from rb = RN(1/b): the same Markstein step (q = RN(a rb)
 * is within one ulp of a/b, the FMA residual is exact, one correction rounds
 * correctly), for quotients that are neither subnormal nor overflowing.  A
 * zero b (rb infinite) gives NaN where a/b would give +-inf or NaN. */
__device__ __forceinline__ double div_cr(double a, double b, double rb)
{
    const double q = a * rb;
    const double r = __builtin_fma(-q, b, a);
    return __builtin_fma(r, rb, q);
}

/* Correctly rounded double square root of 0 or a positive normal float:
 * reciprocal square-root seed and the Goldschmidt/Newton correction used by
 * the AMDGPU f64 sqrt expansion, without its denormal scaling (inputs are
 * >= 2^-149 or exactly 0). */
__device__ __forceinline__ double sqrt_cr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? 0.0 : g;
}

#endif
