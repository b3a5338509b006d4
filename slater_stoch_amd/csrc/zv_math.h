/*
 * Deterministic FP64 math shared by the HIP kernels and host code.
 *
 * The fixed-sample mode promises bit-identical e_S / e_G on the GPU and on
 * a CPU build of the same formulas.  libm's exp (glibc) and the device
 * library's exp differ in the last bit for some arguments, so both sides
 * use this one exp, evaluated with the same sequence of IEEE operations:
 *
 *   x = k ln2 + r,  k = rint(x log2 e),  |r| <= ln2/2
 *   r  = (x - k*ln2_hi) - k*ln2_lo          (Cody-Waite two-constant split)
 *   e^r = sum_{n=0}^{13} r^n / n!           (Horner with fused multiply-add)
 *   e^x = ldexp(e^r, k)
 *
 * The truncation error of the degree-13 Taylor sum on |r| <= 0.35 is
 * below 5e-18, so the result is within about one ulp of the exact value.
 * ln2_hi carries 32 significant bits, so k*ln2_hi is exact for |k| < 2^21.
 * Arguments below -708 return 0 (no subnormal results, so the device and
 * host agree bitwise there too).
 *
 * Every translation unit that includes this header must be compiled with
 * -ffp-contract=off so that no multiply-add is fused behind our back.
 */
#ifndef ZV_MATH_H
#define ZV_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define ZV_HD __host__ __device__ inline
#else
#define ZV_HD static inline
#endif

#define ZV_LN2_HI 6.93147180369123816490e-01
#define ZV_LN2_LO 1.90821492927058770002e-10
#define ZV_LOG2E 1.44269504088896338700e+00

ZV_HD double zv_exp(double x)
{
    if (x != x) return x;
    if (x < -708.0) return 0.0;
    if (x > 709.78) return HUGE_VAL;
    double kd = rint(x * ZV_LOG2E);
    double r = (x - kd * ZV_LN2_HI) - kd * ZV_LN2_LO;
    double p = 1.0 / 6227020800.0; /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kd);
}

#endif /* ZV_MATH_H */
