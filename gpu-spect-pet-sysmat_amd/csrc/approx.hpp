// Branch-free FP32 approximations for the scatter site term, from published
// methods, host and device (the host versions exist for the accuracy tests):
//
//  * erfc_nr(z), z >= 0: the Chebyshev-fitted form of Numerical Recipes in C
//    (2nd ed., section 6.2, erfcc): erfc(z) = t exp(-z^2 + c(t)), t = 1 / (1 + z/2),
//    c a degree-9 polynomial, fractional error below 1.2e-7 for every z >= 0.  The
//    exponent is summed in FP64 (z^2 split exactly into hi + lo), so the large
//    arguments of the tails lose no digits; exp is 2^x with the integer part
//    applied exactly.  Measured: relative error <= 4e-7 against the FP64 erfc on
//    z in [0, 10].
//  * atan2_upper(y, x), y >= 0: octant reduction to a = min / max in [0, 1] and
//    the minimax polynomial of Abramowitz & Stegun 4.4.49 (|error| <= 2e-8 rad
//    on [0, 1]); measured |error| <= 3e-7 rad in FP32 against the FP64 atan2.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define APPROX_HD __host__ __device__
#else
#define APPROX_HD
#endif

namespace sysmat {

APPROX_HD inline float approx_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// 2^y for an FP64 exponent: the integer part exactly (ldexp), the fraction in FP32
APPROX_HD inline float approx_exp2_64(double y) {
    const double yi = std::floor(y);
    const float f = (float)(y - yi);
#ifdef __HIP_DEVICE_COMPILE__
    const float m = __builtin_amdgcn_exp2f(f);
#else
    const float m = std::exp2(f);
#endif
    return std::ldexp(m, (int)yi);
}

APPROX_HD inline float erfc_nr(float z) {
    const float t = approx_rcp(1.0f + 0.5f * z);
    float c = 0.17087277f;
    c = std::fma(c, t, -0.82215223f);
    c = std::fma(c, t, 1.48851587f);
    c = std::fma(c, t, -1.13520398f);
    c = std::fma(c, t, 0.27886807f);
    c = std::fma(c, t, -0.18628806f);
    c = std::fma(c, t, 0.09678418f);
    c = std::fma(c, t, 0.37409196f);
    c = std::fma(c, t, 1.00002368f);
    c = std::fma(c, t, -1.26551223f);
    const float hi = z * z, lo = std::fma(z, z, -hi);  // z^2 = hi + lo exactly
    const double e = ((double)c - (double)lo) - (double)hi;
    // exp(e) = 2^(e log2 e); below 2^-149 the result is 0 (ldexp underflows)
    return e > -104.0 ? t * approx_exp2_64(e * 1.4426950408889634) : 0.0f;
}

APPROX_HD inline float atan2_upper(float y, float x) {
    const float ax = std::fabs(x);
    const bool swap = y > ax;
    const float num = swap ? ax : y, den = swap ? y : ax;
    const float a = den > 0.0f ? num * approx_rcp(den) : 0.0f;
    const float a2 = a * a;
    float p = 0.0028662257f;
    p = std::fma(p, a2, -0.0161657367f);
    p = std::fma(p, a2, 0.0429096138f);
    p = std::fma(p, a2, -0.0752896400f);
    p = std::fma(p, a2, 0.1065626393f);
    p = std::fma(p, a2, -0.1420889944f);
    p = std::fma(p, a2, 0.1999355085f);
    p = std::fma(p, a2, -0.3333314528f);
    float r = std::fma(a * a2, p, a);  // a (1 + a^2 p)
    r = swap ? 1.57079632679489662f - r : r;
    return x < 0.0f ? 3.14159265358979324f - r : r;
}

}  // namespace sysmat
