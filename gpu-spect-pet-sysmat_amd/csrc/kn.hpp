// Klein-Nishina partial integrals without the per-step loop.
//
// The reference evaluates the angular weight of a scatter site as a left
// Riemann sum with step h = 0.01 rad (computeComptonIntegral, scatter.cu:979-993):
//     S(a, n) = h * sum_{i=0..n} f(a + i h),  n = floor((b0 - a) / h),
//     f(theta) = KN(cos theta) * sin theta,
//     KN(u) = P^3 + P - P^2 + u^2 P^2,  P = 1 / (1 + alpha (1 - u)),  alpha = E/511
// (the reference's f2/(1+f1)^2 (1 + f1^2/(f2 (1+f1))) written in P; both equal
// P^2 (P + 1/P - sin^2 theta)).  Here the sum is evaluated in O(1) with the
// Euler-Maclaurin formula on the exact antiderivative:
//     S = int_a^b f + h (f(a) + f(b))/2 + h^2 (f'(b) - f'(a))/12 + O(h^4 f'''),  b = a + n h,
//     int_a^b f dtheta = (1/alpha) [H(w(b)) - H(w(a))],  w = 1 + alpha (1 - cos theta),
//     H(w) = -1/(2 w^2) + ln w + 1/w + (w - beta^2/w - 2 beta ln w) / alpha^2,  beta = 1 + alpha.
// The neglected term is below 1e-10 of S for h = 0.01; the step count n is
// taken exactly as the reference does, so the discrete structure is kept.
// The caller passes cos / sin of the start angle (the scatter kernel has cos
// theta_min itself) and the tabulated (cos n h, sin n h), so no trigonometric
// function is evaluated per site.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define KN_HD __host__ __device__
#else
#define KN_HD
#endif

namespace sysmat {

struct KnCoef {
    double alpha, beta, ia, ia2;  // E/511, 1 + alpha, 1/alpha, 1/alpha^2
};

// 1 / x for x >= 1: on the GPU the FP64 reciprocal with two Newton steps (~1 ulp,
// against the ~10 instructions of the correctly rounded division), on the host the
// division.
KN_HD inline double kn_rcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
#else
    return 1.0 / x;
#endif
}

KN_HD inline KnCoef kn_coef(double e0) {
    KnCoef c;
    c.alpha = e0 / 511.0;
    c.beta = 1.0 + c.alpha;
    c.ia = 1.0 / c.alpha;
    c.ia2 = c.ia * c.ia;
    return c;
}

// f(theta) and f'(theta) at cos = u, sin = s, with P = 1 / (1 + alpha (1 - u)).
KN_HD inline void kn_fP(const KnCoef& c, double u, double s, double P, double& f, double& df) {
    const double P2 = P * P, P3 = P2 * P;
    const double g = P3 + P - P2 + u * u * P2;
    const double dg = c.alpha * (3.0 * P2 * P2 + P2 - 2.0 * P3) + 2.0 * u * P2 + 2.0 * c.alpha * u * u * P3;
    f = g * s;
    df = -dg * s * s + g * u;
}

// ln(1 + x) = 2 atanh(s), s = x / (2 + x), as 2 s (1 + s^2/3 + ... + s^32/33): within
// ~2 ulp for s^2 <= 0.1, against ~135 instructions of the library's FP64 log1p.  The KN
// argument x = (w_b - w_a) / w_a has |s| <= alpha / (1 + alpha), so the series covers
// every argument when alpha = E/511 <= kKnSeriesMaxAlpha (E <= 236 keV).
KN_HD inline double log1p_series(double x) {
    const double s = x * kn_rcp(2.0 + x);
    const double s2 = s * s;
    double p = 1.0 / 33.0;
#pragma unroll
    for (int k = 15; k >= 0; --k) p = fma(p, s2, 1.0 / (double)(2 * k + 1));
    return 2.0 * s * p;
}
constexpr double kKnSeriesMaxAlpha = 0.462;  // alpha / (1 + alpha) <= sqrt(0.1)

// h * sum_{i=0..n} f(a + i h) from cos / sin of a and (cos n h, sin n h): the end
// angle b = a + n h by angle addition, the antiderivative difference through one
// log1p,
//   H(wb) - H(wa) = -(1/wb^2 - 1/wa^2)/2 + ln(wb/wa) + (1/wb - 1/wa)
//                   + ((wb - wa) - beta^2 (1/wb - 1/wa) - 2 beta ln(wb/wa)) / alpha^2,
// with wb - wa = alpha (cos a - cos b) taken directly (no cancellation).
// kSeries: ln(w_b / w_a) by log1p_series (only for alpha <= kKnSeriesMaxAlpha).
template <bool kSeries = false>
KN_HD inline double kn_sum_cs(const KnCoef& c, double ua, double sa, double cnh, double snh, double h) {
    const double ub = ua * cnh - sa * snh, sb = sa * cnh + ua * snh;
    const double wa = 1.0 + c.alpha * (1.0 - ua), wb = 1.0 + c.alpha * (1.0 - ub);
    const double iab = kn_rcp(wa * wb);  // one reciprocal for both
    const double iwa = wb * iab, iwb = wa * iab;
    double fa, dfa, fb, dfb;
    kn_fP(c, ua, sa, iwa, fa, dfa);
    kn_fP(c, ub, sb, iwb, fb, dfb);
    const double dw = c.alpha * (ua - ub);
    const double diw = -dw * iwa * iwb;
    const double dl = kSeries ? log1p_series(dw * iwa) : log1p(dw * iwa);
    const double dH = -0.5 * diw * (iwb + iwa) + dl + diw + (dw - c.beta * c.beta * diw - 2.0 * c.beta * dl) * c.ia2;
    return dH * c.ia + 0.5 * h * (fa + fb) + h * h / 12.0 * (dfb - dfa);
}

}  // namespace sysmat
