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

KN_HD inline KnCoef kn_coef(double e0) {
    KnCoef c;
    c.alpha = e0 / 511.0;
    c.beta = 1.0 + c.alpha;
    c.ia = 1.0 / c.alpha;
    c.ia2 = c.ia * c.ia;
    return c;
}

KN_HD inline double kn_H(const KnCoef& c, double w) {
    const double lw = log(w), iw = 1.0 / w;
    return -0.5 * iw * iw + lw + iw + (w - c.beta * c.beta * iw - 2.0 * c.beta * lw) * c.ia2;
}

// f(theta) and f'(theta) at cos = u, sin = s.
KN_HD inline void kn_f(const KnCoef& c, double u, double s, double& f, double& df) {
    const double P = 1.0 / (1.0 + c.alpha * (1.0 - u));
    const double P2 = P * P, P3 = P2 * P;
    const double g = P3 + P - P2 + u * u * P2;
    const double dg = c.alpha * (3.0 * P2 * P2 + P2 - 2.0 * P3) + 2.0 * u * P2 + 2.0 * c.alpha * u * u * P3;
    f = g * s;
    df = -dg * s * s + g * u;
}

// h * sum_{i=0..n} f(a + i h)
KN_HD inline double kn_sum(const KnCoef& c, double a, int n, double h) {
    const double b = a + (double)n * h;
    const double ua = cos(a), sa = sin(a), ub = cos(b), sb = sin(b);
    double fa, dfa, fb, dfb;
    kn_f(c, ua, sa, fa, dfa);
    kn_f(c, ub, sb, fb, dfb);
    const double wa = 1.0 + c.alpha * (1.0 - ua), wb = 1.0 + c.alpha * (1.0 - ub);
    const double integral = (kn_H(c, wb) - kn_H(c, wa)) * c.ia;
    return integral + 0.5 * h * (fa + fb) + h * h / 12.0 * (dfb - dfa);
}

}  // namespace sysmat
